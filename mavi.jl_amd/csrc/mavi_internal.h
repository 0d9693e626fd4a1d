// mavi_internal.h -- shared types and device helpers of libmavi_hip.so.
//
// Device data layout (per system, all in HBM, allocated once at creation):
//   pos, vel, force        double2[n]      id order (slot order of state.pos)
//   key, rank              int32[n]        cell of each id / arrival rank in it
//   cnt, start             int32[ncell+1]  particles per cell / exclusive scan
//   perm                   int32[na]       sorted slot -> id  (cells row-major,
//                                          ids ascending inside a cell)
//   scell                  int32[na]       cell of each sorted slot
//   ps, ps2                double2[na]     sorted positions (x_n, drifted x')
//   f1                     double2[na]     first-evaluation force, sorted order
// where na = number of active ids (get_particles_ids) and the cell index is
// row0*cols + col0 with row0 = reference row - 1 (row 1 = top of the box,
// src/chunks.jl:129) and col0 = reference col - 1.  Row-major cells make the
// three cells (row, col-1..col+1) one contiguous range of sorted slots.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/mavi_hip.h"

#define MAVI_FLAG_OVERFLOW 1u
#define MAVI_FLAG_OUT_OF_BOX 2u
#define MAVI_FLAG_NONFINITE 4u

struct mavi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
};

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
    double lj24eps, ljsig2;
    // rings
    int32_t nmax, nrings, ntypes;
    int32_t typed;
};

// Per ring type parameters (device constant-like array).
struct RingType {
    double p0, relax, vo, mu, rot, karea, kspring, lspring;
    int32_t np;
    int32_t pad;
};

// Interaction (HarmTrunc) for one pair of ring types, with exact thresholds.
struct RingInter {
    double k_rep, k_atr, dist_eq, dist_max;
    double max2;   // dist > dist_max <=> r2 > max2
    double eq2;    // dist < dist_eq  <=> r2 < eq2
};

struct ProfSlot {
    double ms = 0.0;
    int64_t launches = 0;
};

struct mavi_sys {
    mavi_ctx *ctx = nullptr;
    SysParams P{};
    int32_t wg = 256;  // pair-force workgroup size
    // device buffers
    double2 *pos = nullptr, *vel = nullptr, *force = nullptr;
    int32_t *key = nullptr, *rank = nullptr, *cnt = nullptr, *start = nullptr;
    int32_t *bsum = nullptr;  // scan block sums
    int32_t *perm = nullptr, *scell = nullptr;
    double2 *ps = nullptr, *ps2 = nullptr, *f1 = nullptr;
    uint32_t *flags = nullptr;
    int32_t *active_ids = nullptr;  // na ids ascending (rings with padding)
    // rings
    double *pol = nullptr, *areas = nullptr;
    double2 *cms = nullptr, *cms_next = nullptr, *cpos = nullptr;
    int32_t *rtype = nullptr;   // per ring, 0-based
    RingType *types = nullptr;  // ntypes
    RingInter *inter = nullptr; // ntypes*ntypes
    double *noise = nullptr;    // host-fed noise staging
    int64_t noise_cap = 0;
    uint64_t seed = 0;
    uint64_t rng_step = 0;      // device noise counter (steps taken)
    // observables scratch
    double *red = nullptr;
    int32_t red_cap = 0;
    int32_t nscan_blocks = 0;
    bool has_state = false;
    bool binned = false;        // cells valid for current positions
    // graph of G steps
    hipGraphExec_t graph = nullptr;
    int64_t graph_steps = 0;
    // profiling
    bool prof = false;
    ProfSlot profs[MAVI_PROF_NCLASS];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// --------------------------------------------------------------------------
// Device helpers
// --------------------------------------------------------------------------

// Exact truncated quotient of a/b, equal to Julia's Base div(a, b) for
// Float64 (src/chunks.jl:129-130): q = trunc(a/b) can only be off by one when
// the rounded quotient lands on an integer; the FMA residual a - q*b is exact
// in sign and detects that case.
__device__ __forceinline__ double exact_trunc_div(double a, double b) {
    double q = trunc(a / b);
    double r = __builtin_fma(-q, b, a);
    if (a >= 0.0) {
        if (r < 0.0) q -= 1.0;
    } else {
        if (r > 0.0) q += 1.0;
    }
    return q;
}

// calc_diff for a periodic rectangle, src/integration.jl:43-48.
__device__ __forceinline__ double min_image(double dr, double size, double half) {
    double m = (fabs(dr) > half) ? 1.0 : 0.0;
    return dr - m * copysign(size, dr);
}

// Host-side launchers (mavi_kernels.hip)
namespace mavi {
hipError_t launch_bin(mavi_sys *s, hipStream_t st);
hipError_t launch_export_chunks(mavi_sys *s, int64_t *d_cp, int64_t *d_counts, hipStream_t st);
hipError_t launch_forces_only(mavi_sys *s, hipStream_t st);
hipError_t launch_verlet_step(mavi_sys *s, hipStream_t st);
hipError_t launch_rings_step(mavi_sys *s, const double *noise_dev, uint64_t step_index, hipStream_t st);
hipError_t launch_rings_forces(mavi_sys *s, hipStream_t st);
hipError_t launch_rings_info(mavi_sys *s, hipStream_t st);
hipError_t launch_observables(mavi_sys *s, double *d_out, hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);
}  // namespace mavi

// Host helpers (mavi_capi.hip)
double exact_gt_threshold(double rc);   // largest r2 with sqrt(r2) <= rc
double exact_lt_threshold(double deq);  // smallest r2 with sqrt(r2) >= deq
