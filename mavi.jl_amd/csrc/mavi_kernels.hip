// mavi_kernels.hip -- gfx950 kernels of the step hot path.
//
//   K1 binning      k_bin_count -> k_scan_* -> k_scatter -> k_cell_finalize
//                   (counting sort by chunk; ids ascending inside a chunk, as
//                   update_chunks! visits them, src/chunks.jl:150-161)
//   K2 pair forces  k_pair<MODE, POT>: gather over the 3x3 chunk block around
//                   each particle; the stencil of src/chunks.jl:61-118 visits
//                   each adjacent chunk pair once, the gather visits it from
//                   both sides, so the interaction sets are equal for grids
//                   >= 3x3 (src/integration.jl:112-157).
//   K3 Verlet+walls MODE_VERLET1/2 epilogues of k_pair (src/integration.jl:
//                   415-431 with walls! :271-285 / :309-324).
//   K4 rings        k_rings<UPDATE>: per-ring unwrap, shoelace area,
//                   springs, area force and overdamped update
//                   (src/rings/integration.jl:79-351).
//   K7 observables  k_observables (src/quantities.jl:12-18, 46-66).
//
// The library is compiled with -ffp-contract=off: a*b+c is only fused where
// the code calls __builtin_fma, so the arithmetic that must follow Julia's
// rounding (positions, velocities, HarmTrunc forces, ring updates) does.
#include "mavi_internal.h"

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
    // v_rcp_f64 followed by two Newton steps: |error| ~ 1 ulp, enough for
    // the 1e-12 force tolerance of the LJ path.
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    return y;
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// K1 binning
// ---------------------------------------------------------------------------

// Cell of one position, update_particle_chunk!, src/chunks.jl:120-142.
__device__ __forceinline__ int32_t cell_of(const SysParams &P, double x, double y, uint32_t *flag) {
    double qr = exact_trunc_div((-y + P.bly) + P.H, P.ch);
    double qc = exact_trunc_div(x - P.blx, P.cl);
    if (!(fabs(qr) < 1e15) || !(fabs(qc) < 1e15)) {  // NaN / Inf / absurd
        *flag |= MAVI_FLAG_NONFINITE;
        return 0;
    }
    int64_t r = (int64_t)qr;  // 0-based (reference row - 1)
    int64_t c = (int64_t)qc;
    r -= (r == P.rows) ? 1 : 0;  // the N+1 -> N clamp
    c -= (c == P.cols) ? 1 : 0;
    if (r < 0 || r >= P.rows || c < 0 || c >= P.cols) {
        *flag |= MAVI_FLAG_OUT_OF_BOX;
        r = r < 0 ? 0 : (r >= P.rows ? P.rows - 1 : r);  // keep memory access in bounds
        c = c < 0 ? 0 : (c >= P.cols ? P.cols - 1 : c);
    }
    return (int32_t)(r * P.cols + c);
}

__global__ void k_bin_count(SysParams P, const double2 *__restrict__ pos, const int32_t *__restrict__ active,
                            int32_t *__restrict__ key, int32_t *__restrict__ rank, int32_t *__restrict__ cnt,
                            uint32_t *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.na) return;
    int id = active ? active[i] : i;
    double2 p = pos[id];
    uint32_t flag = 0;
    int32_t c = cell_of(P, p.x, p.y, &flag);
    if (flag) atomicOr(flags, flag);
    key[i] = c;
    rank[i] = atomicAdd(&cnt[c], 1);
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

__global__ void k_scan_local(const int32_t *__restrict__ in, int32_t *__restrict__ out, int32_t *__restrict__ bsum,
                             int32_t n) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int32_t total;
    int32_t off = block_excl_scan(s, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the block sums in one workgroup, tile by tile.
__global__ void k_scan_bsum(int32_t *__restrict__ bsum, int32_t nb) {
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int t0 = 0; t0 < nb; t0 += kScanThreads) {
        int i = t0 + threadIdx.x;
        int32_t v = (i < nb) ? bsum[i] : 0;
        int32_t total;
        int32_t ex = block_excl_scan(v, &total);
        int32_t c = carry;
        if (i < nb) bsum[i] = ex + c;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
}

__global__ void k_scan_add(int32_t *__restrict__ out, const int32_t *__restrict__ bsum, int32_t n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / kScanTile];
}

__global__ void k_scatter(SysParams P, const int32_t *__restrict__ active, const int32_t *__restrict__ key,
                          const int32_t *__restrict__ rank, const int32_t *__restrict__ start,
                          int32_t *__restrict__ perm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.na) return;
    int id = active ? active[i] : i;
    perm[start[key[i]] + rank[i]] = id;
}

// One thread per chunk: restore ascending id order inside the chunk (the
// order update_chunks! appends them in), check the capacity nc, and gather
// the chunk's positions into sorted order.
__global__ void k_cell_finalize(SysParams P, const int32_t *__restrict__ start, int32_t *__restrict__ perm,
                                const double2 *__restrict__ pos, double2 *__restrict__ ps,
                                int32_t *__restrict__ scell, uint32_t *__restrict__ flags) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    const int b = start[c], e = start[c + 1], m = e - b;
    if (m == 0) return;
    if (m > P.nc) atomicOr(flags, MAVI_FLAG_OVERFLOW);
    constexpr int R = 16;
    if (m <= R) {
        int32_t v[R];
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = (k < m) ? perm[b + k] : 0x7fffffff;
        // insertion sort in registers (fully unrolled, fixed network)
#pragma unroll
        for (int k = 1; k < R; k++) {
#pragma unroll
            for (int q = k; q > 0; q--) {
                int32_t lo = min(v[q - 1], v[q]), hi = max(v[q - 1], v[q]);
                v[q - 1] = lo;
                v[q] = hi;
            }
        }
#pragma unroll
        for (int k = 0; k < R; k++)
            if (k < m) {
                perm[b + k] = v[k];
                ps[b + k] = pos[v[k]];
                scell[b + k] = c;
            }
    } else {
        for (int k = b + 1; k < e; k++) {
            int32_t x = perm[k];
            int q = k - 1;
            while (q >= b && perm[q] > x) {
                perm[q + 1] = perm[q];
                q--;
            }
            perm[q + 1] = x;
        }
        for (int k = b; k < e; k++) {
            ps[k] = pos[perm[k]];
            scell[k] = c;
        }
    }
}

// chunk_particles (nc, rows, cols) and num_particles_in_chunk (rows, cols),
// column-major, 1-based: the reference layout of src/chunks.jl:23-24.
__global__ void k_export_chunks(SysParams P, const int32_t *__restrict__ start, const int32_t *__restrict__ perm,
                                int64_t *__restrict__ cp, int64_t *__restrict__ counts) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    int row0 = c / P.cols, col0 = c - row0 * P.cols;
    int64_t k = (int64_t)col0 * P.rows + row0;
    int b = start[c], m = start[c + 1] - b;
    counts[k] = m;
    int lim = m < P.nc ? m : P.nc;
    for (int q = 0; q < lim; q++) cp[(int64_t)P.nc * k + q] = (int64_t)perm[b + q] + 1;
}

// Gather for chunks = nothing: sorted order is the active id order.
__global__ void k_gather_active(SysParams P, const int32_t *__restrict__ active, const double2 *__restrict__ pos,
                                int32_t *__restrict__ perm, double2 *__restrict__ ps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.na) return;
    int id = active ? active[i] : i;
    perm[i] = id;
    ps[i] = pos[id];
}

// ---------------------------------------------------------------------------
// K2/K3 pair forces with fused integrator epilogues
// ---------------------------------------------------------------------------

enum { MODE_FORCES = 0, MODE_VERLET1 = 1, MODE_VERLET2 = 2 };
enum { POT_LJ = 0, POT_HT = 1, POT_RINGS = 2 };

struct RingCtx {
    const int32_t *perm;
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

template <int POT>
struct PairAcc {
    double fx = 0.0, fy = 0.0;
    // pair context for rings
    int32_t id_i = 0, ring_i = 0, type_i = 0;

    __device__ __forceinline__ void add(const SysParams &P, const RingCtx &R, double dx, double dy, bool self_ok,
                                        int j) {
        const double r2 = dx * dx + dy * dy;  // sum(dr.^2): two roundings, no FMA
        if constexpr (POT == POT_LJ) {
            const bool ok = self_ok && (r2 <= P.cut2);
            const double inv = rcp_nr(r2);
            const double s2 = P.ljsig2 * inv;
            const double s6 = s2 * s2 * s2;
            double fr = P.lj24eps * s6 * __builtin_fma(2.0, s6, -1.0) * inv;
            fr = ok ? fr : 0.0;
            fx = __builtin_fma(fr, dx, fx);
            fy = __builtin_fma(fr, dy, fy);
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
            const int32_t id_j = R.perm[j];
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

template <int POT>
__device__ __forceinline__ void scan_range(const SysParams &P, const RingCtx &R, PairAcc<POT> &acc,
                                           const double2 *__restrict__ pin, int s, double xi, double yi,
                                           bool wrap, int a, int b) {
    for (int j = a; j < b; j++) {
        const double2 pj = pin[j];
        double dx = xi - pj.x, dy = yi - pj.y;
        if (wrap) {
            dx = min_image(dx, P.L, P.halfL);
            dy = min_image(dy, P.H, P.halfH);
        }
        acc.add(P, R, dx, dy, j != s, j);
    }
}

template <int MODE, int POT, bool ALLPAIRS>
__global__ void k_pair(SysParams P, RingCtx R, const double2 *__restrict__ pin, const int32_t *__restrict__ scell,
                       const int32_t *__restrict__ start, const int32_t *__restrict__ perm,
                       double2 *__restrict__ pos, double2 *__restrict__ vel, double2 *__restrict__ force,
                       double2 *__restrict__ ps2, double2 *__restrict__ f1) {
    const int s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= P.na) return;
    const double2 pi = pin[s];
    PairAcc<POT> acc;
    if constexpr (POT == POT_RINGS) {
        acc.id_i = perm[s];
        acc.ring_i = ring_of((uint32_t)acc.id_i, R);
        acc.type_i = P.typed ? R.rtype[acc.ring_i] : 0;
    }
    if constexpr (ALLPAIRS) {
        scan_range<POT>(P, R, acc, pin, s, pi.x, pi.y, P.periodic != 0, 0, P.na);
    } else {
        const int c = scell[s];
        const int row = c / P.cols, col = c - row * P.cols;
        const bool wrap = P.periodic && (row == 0 || row == P.rows - 1 || col == 0 || col == P.cols - 1);
        for (int d = -1; d <= 1; d++) {
            int r = row + d;
            if (P.periodic) r = (r < 0) ? r + P.rows : (r >= P.rows ? r - P.rows : r);
            else if (r < 0 || r >= P.rows) continue;
            const int base = r * P.cols;
            int a0, a1, b0 = 0, b1 = 0;
            if (!P.periodic) {
                const int c0 = col > 0 ? col - 1 : 0, c1 = col < P.cols - 1 ? col + 1 : P.cols - 1;
                a0 = start[base + c0];
                a1 = start[base + c1 + 1];
            } else if (col == 0) {
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
    const double fx = acc.fx, fy = acc.fy;
    if constexpr (MODE == MODE_FORCES) {
        force[perm[s]] = make_double2(fx, fy);
    } else if constexpr (MODE == MODE_VERLET1) {
        // update_verlet!: pos += vel*dt + forces*term (src/integration.jl:424)
        const int id = perm[s];
        const double2 v = vel[id];
        ps2[s] = make_double2(pi.x + (v.x * P.dt + fx * P.term), pi.y + (v.y * P.dt + fy * P.term));
        f1[s] = make_double2(fx, fy);
    } else {
        // vel += dt/2 * (forces + old_forces) (src/integration.jl:430), then walls!
        const int id = perm[s];
        const double2 v = vel[id];
        const double2 fo = f1[s];
        double vx = v.x + P.hdt * (fx + fo.x);
        double vy = v.y + P.hdt * (fy + fo.y);
        double x = pi.x, y = pi.y;
        if (P.periodic) {
            // src/integration.jl:309-324
            const double dxc = x - (P.blx + P.halfL), dyc = y - (P.bly + P.halfH);
            const bool ox = fabs(dxc) > P.halfL, oy = fabs(dyc) > P.halfH;
            if (ox || oy) {
                if (ox) x = x - copysign(P.halfL * 2.0, dxc);
                if (oy) y = y - copysign(P.halfH * 2.0, dyc);
            }
        } else {
            // src/integration.jl:271-285
            const double rx = x - P.blx, ry = y - P.bly;
            const bool ox = ((rx + P.radius) > P.L) || ((rx - P.radius) < 0.0);
            const bool oy = ((ry + P.radius) > P.H) || ((ry - P.radius) < 0.0);
            if (ox) vx = -vx;
            if (oy) vy = -vy;
        }
        pos[id] = make_double2(x, y);
        vel[id] = make_double2(vx, vy);
        force[id] = make_double2(fx, fy);
    }
}

// ---------------------------------------------------------------------------
// K4 rings: one thread per ring
// ---------------------------------------------------------------------------

__device__ __forceinline__ double2 cdiff(const SysParams &P, double2 a, double2 b) {
    double dx = a.x - b.x, dy = a.y - b.y;
    if (P.periodic) {
        dx = min_image(dx, P.L, P.halfL);
        dy = min_image(dy, P.H, P.halfH);
    }
    return make_double2(dx, dy);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -> one standard normal
// through Box-Muller on two 53-bit uniforms.
__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                             uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
    uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
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
    const double u1 = ((double)(((uint64_t)c0 << 21) ^ (uint64_t)c1 >> 11) + 1.0) * 0x1p-53;  // (0,1]
    const double u2 = (double)((((uint64_t)c2 << 21) ^ ((uint64_t)c3 >> 11)) & ((1ull << 53) - 1)) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

template <bool UPDATE>
__global__ void k_rings(SysParams P, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                        double2 *__restrict__ pos, double2 *__restrict__ force, double *__restrict__ pol,
                        double *__restrict__ areas, double2 *__restrict__ cms, double2 *__restrict__ cms_next,
                        double2 *__restrict__ cpos, const double *__restrict__ noise, uint64_t seed,
                        uint64_t step_index) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings) return;
    const RingType T = types[P.typed ? rtype[r] : 0];
    const int np = T.np;
    const int64_t base = (int64_t)r * P.nmax;
    double2 *x = pos + base;
    double2 *F = force + base;
    double2 *c = cpos + base;

    // update_cms! (src/rings/integration.jl:366-372) reads the continuous
    // positions of the previous update_continuos_pos!
    if constexpr (UPDATE) cms[r] = cms_next[r];

    // update_continuos_pos!, :118-136 (copies every slot, unwraps the ring)
    for (int i = 0; i < P.nmax; i++) c[i] = x[i];
    if (P.periodic) {
        for (int i = 1; i < np; i++) {
            const double2 d = cdiff(P, x[i], x[i - 1]);
            c[i] = make_double2(c[i - 1].x + d.x, c[i - 1].y + d.y);
        }
    }
    {
        double sx = 0.0, sy = 0.0;
        for (int i = 0; i < np; i++) {
            sx += c[i].x;
            sy += c[i].y;
        }
        cms_next[r] = make_double2(sx / (double)np, sy / (double)np);
    }
    // calc_area, :103-116
    double area = 0.0;
    for (int i = 0; i < np - 1; i++) area += c[i].x * c[i + 1].y - c[i].y * c[i + 1].x;
    area += c[np - 1].x * c[0].y - c[np - 1].y * c[0].x;
    area = area / 2.0;
    areas[r] = area;

    // springs (forces!, :205-223) in the reference's accumulation order:
    // particle 1 gets +f_1 then -f_n, particle i > 1 gets -f_{i-1} then +f_i.
    auto spring = [&](int a, int b) {
        const double2 d = cdiff(P, x[a], x[b]);
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
        double2 Fi = F[i];
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
        F[i] = Fi;
    }
    // area_forces!, :157-194 (after all springs, as in the reference)
    for (int i = 0; i < np; i++) {
        const int i1 = (i == 0) ? np - 1 : i - 1;
        const int i2 = (i == np - 1) ? 0 : i + 1;
        const double2 d = cdiff(P, x[i2], x[i1]);
        double2 Fi = F[i];
        Fi.x -= fa * (d.y / 2.0);
        Fi.y -= fa * (-d.x / 2.0);
        F[i] = Fi;
    }
    if constexpr (!UPDATE) return;

    // update!, :300-351
    const double theta = pol[r];
    double sn, cs;
    sincos(theta, &sn, &cs);
    double vcx = 0.0, vcy = 0.0;
    for (int i = 0; i < np; i++) {
        const double2 Fi = F[i];
        const double vx = T.vo * cs + T.mu * Fi.x;
        const double vy = T.vo * sn + T.mu * Fi.y;
        vcx += vx;
        vcy += vy;
        double px = x[i].x + vx * P.dt, py = x[i].y + vy * P.dt;
        if (P.periodic) {  // walls! after update!, applied per particle
            const double dxc = px - (P.blx + P.halfL), dyc = py - (P.bly + P.halfH);
            const bool ox = fabs(dxc) > P.halfL, oy = fabs(dyc) > P.halfH;
            if (ox) px = px - copysign(P.halfL * 2.0, dxc);
            if (oy) py = py - copysign(P.halfH * 2.0, dyc);
        }
        x[i] = make_double2(px, py);
    }
    vcx = vcx / (double)np;
    vcy = vcy / (double)np;
    const double speed = sqrt(vcx * vcx + vcy * vcy);
    double cr = 0.0;
    if (speed != 0.0) {
        cr = (cs * vcy - sn * vcx) / speed;
        if (fabs(cr) > 1.0) cr = cr > 0.0 ? 1.0 : -1.0;
    }
    const double z = noise ? noise[r] : (T.rot != 0.0 ? philox_normal(seed, step_index, (uint32_t)r) : 0.0);
    const double d_theta = 1.0 / T.relax * asin(cr) * P.dt + sqrt(2.0 * T.rot * P.dt) * z;
    pol[r] = theta + d_theta;
}

__global__ void k_rings_init(SysParams P, const int32_t *__restrict__ rtype, const RingType *__restrict__ types,
                             const double2 *__restrict__ pos, double2 *__restrict__ cms,
                             double2 *__restrict__ cms_next, double2 *__restrict__ cpos, double *__restrict__ areas) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.nrings) return;
    const int np = types[P.typed ? rtype[r] : 0].np;
    const int64_t base = (int64_t)r * P.nmax;
    const double2 *x = pos + base;
    double2 *c = cpos + base;
    for (int i = 0; i < P.nmax; i++) c[i] = x[i];
    if (P.periodic)
        for (int i = 1; i < np; i++) {
            const double2 d = cdiff(P, x[i], x[i - 1]);
            c[i] = make_double2(c[i - 1].x + d.x, c[i - 1].y + d.y);
        }
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < np; i++) {
        sx += c[i].x;
        sy += c[i].y;
    }
    cms[r] = cms_next[r] = make_double2(sx / (double)np, sy / (double)np);
    double area = 0.0;
    for (int i = 0; i < np - 1; i++) area += c[i].x * c[i + 1].y - c[i].y * c[i + 1].x;
    area += c[np - 1].x * c[0].y - c[np - 1].y * c[0].x;
    areas[r] = area / 2.0;
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
    if (i < P.n) {
        if (vel) {
            const double2 v = vel[i];
            ke = v.x * v.x + v.y * v.y;
        }
        if (with_pe) {
            const double2 pi = pos[i];
            const double sig = P.p0;
            for (int j = i + 1; j < P.n; j++) {
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

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

RingCtx ring_ctx(mavi_sys *s) {
    RingCtx R{};
    R.perm = s->perm;
    R.rtype = s->rtype;
    R.types = s->types;
    R.inter = s->inter;
    R.nmax = (uint32_t)(s->P.nmax > 0 ? s->P.nmax : 1);
    R.inv_nmax = 1.0f / (float)R.nmax;
    return R;
}

template <int MODE, int POT, bool ALL>
hipError_t launch_pair_t(mavi_sys *s, const double2 *pin, hipStream_t st) {
    const int wg = s->wg;
    hipLaunchKernelGGL((k_pair<MODE, POT, ALL>), dim3(nblk(s->P.na, wg)), dim3(wg), 0, st, s->P, ring_ctx(s), pin,
                       s->scell, s->start, s->perm, s->pos, s->vel, s->force, s->ps2, s->f1);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_pair(mavi_sys *s, const double2 *pin, hipStream_t st) {
    const bool all = s->P.cols == 0;
    switch (s->P.potential) {
        case MAVI_POT_LJ:
            return all ? launch_pair_t<MODE, POT_LJ, true>(s, pin, st) : launch_pair_t<MODE, POT_LJ, false>(s, pin, st);
        case MAVI_POT_HARMTRUNC:
            return all ? launch_pair_t<MODE, POT_HT, true>(s, pin, st) : launch_pair_t<MODE, POT_HT, false>(s, pin, st);
        default:
            return all ? launch_pair_t<MODE, POT_RINGS, true>(s, pin, st)
                       : launch_pair_t<MODE, POT_RINGS, false>(s, pin, st);
    }
}

struct ProfScope {
    mavi_sys *s;
    int which;
    hipStream_t st;
    ProfScope(mavi_sys *s_, int w, hipStream_t st_) : s(s_), which(w), st(st_) {
        if (s->prof) (void)hipEventRecord(s->ev0, st);
    }
    ~ProfScope() {
        if (!s->prof) return;
        (void)hipEventRecord(s->ev1, st);
        if (hipEventSynchronize(s->ev1) == hipSuccess) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) {
                s->profs[which].ms += ms;
                s->profs[which].launches += 1;
            }
        }
    }
};

}  // namespace

namespace mavi {

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_bin(mavi_sys *s, hipStream_t st) {
    const SysParams &P = s->P;
    if (P.cols == 0) {
        hipLaunchKernelGGL(k_gather_active, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->active_ids, s->pos, s->perm,
                           s->ps);
        return hipGetLastError();
    }
    ProfScope ps(s, MAVI_PROF_BIN, st);
    hipError_t e;
    if ((e = launch_fill_i32(s->cnt, 0, (int64_t)P.ncell + 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin_count, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->pos, s->active_ids, s->key,
                       s->rank, s->cnt, s->flags);
    const int n = P.ncell + 1;
    const int nb = nblk(n, kScanTile);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanThreads), 0, st, s->cnt, s->start, s->bsum, n);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(kScanThreads), 0, st, s->bsum, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, st, s->start, s->bsum, n);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(P.na, 256)), dim3(256), 0, st, P, s->active_ids, s->key, s->rank,
                       s->start, s->perm);
    hipLaunchKernelGGL(k_cell_finalize, dim3(nblk(P.ncell, 256)), dim3(256), 0, st, P, s->start, s->perm, s->pos,
                       s->ps, s->scell, s->flags);
    return hipGetLastError();
}

hipError_t launch_export_chunks(mavi_sys *s, int64_t *d_cp, int64_t *d_counts, hipStream_t st) {
    hipLaunchKernelGGL(k_export_chunks, dim3(nblk(s->P.ncell, 256)), dim3(256), 0, st, s->P, s->start, s->perm,
                       d_cp, d_counts);
    return hipGetLastError();
}

hipError_t launch_forces_only(mavi_sys *s, hipStream_t st) {
    ProfScope ps(s, MAVI_PROF_FORCE, st);
    return launch_pair<MODE_FORCES>(s, s->ps, st);
}

hipError_t launch_verlet_step(mavi_sys *s, hipStream_t st) {
    hipError_t e = launch_bin(s, st);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_VERLET1>(s, s->ps, st)) != hipSuccess) return e;
    }
    ProfScope ps(s, MAVI_PROF_FORCE, st);
    return launch_pair<MODE_VERLET2>(s, s->ps2, st);
}

hipError_t launch_rings_forces(mavi_sys *s, hipStream_t st) {
    hipError_t e = launch_bin(s, st);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_FORCES>(s, s->ps, st)) != hipSuccess) return e;
    }
    ProfScope ps(s, MAVI_PROF_RINGS, st);
    hipLaunchKernelGGL(k_rings<false>, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types,
                       s->pos, s->force, s->pol, s->areas, s->cms, s->cms_next, s->cpos, nullptr, s->seed,
                       (uint64_t)0);
    return hipGetLastError();
}

hipError_t launch_rings_step(mavi_sys *s, const double *noise_dev, uint64_t step_index, hipStream_t st) {
    hipError_t e = launch_bin(s, st);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(s, MAVI_PROF_FORCE, st);
        if ((e = launch_pair<MODE_FORCES>(s, s->ps, st)) != hipSuccess) return e;
    }
    ProfScope ps(s, MAVI_PROF_RINGS, st);
    hipLaunchKernelGGL(k_rings<true>, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types, s->pos,
                       s->force, s->pol, s->areas, s->cms, s->cms_next, s->cpos, noise_dev, s->seed, step_index);
    return hipGetLastError();
}

hipError_t launch_rings_info(mavi_sys *s, hipStream_t st) {
    hipLaunchKernelGGL(k_rings_init, dim3(nblk(s->P.nrings, 64)), dim3(64), 0, st, s->P, s->rtype, s->types, s->pos,
                       s->cms, s->cms_next, s->cpos, s->areas);
    return hipGetLastError();
}

hipError_t launch_observables(mavi_sys *s, double *d_out, hipStream_t st) {
    const int with_pe = (s->P.potential == MAVI_POT_LJ) ? 1 : 0;
    hipLaunchKernelGGL(k_observables, dim3(nblk(s->P.n, kObsThreads)), dim3(kObsThreads), 0, st, s->P, s->pos,
                       s->vel, with_pe, d_out);
    return hipGetLastError();
}

}  // namespace mavi
