// mavi_capi.hip -- the extern "C" boundary of libmavi_hip.so (include/mavi_hip.h).
#include <math.h>
#include <string.h>

#include <cstddef>
#include <cstdio>
#include <string>
#include <algorithm>
#include <vector>

#include "mavi_host.h"
#include "mavi_philox.h"

namespace {

// Self-propelled system (SelfPropelledState with SzaboCfg / RunTumbleCfg).
bool is_active(const mavi_sys *s) {
    return s->P.potential == MAVI_POT_SZABO || s->P.potential == MAVI_POT_RTP;
}

// Turn device flags raised since the last check into a status.
int check_flags(mavi_sys *s) {
    mavi_ctx *ctx = s->ctx;
    uint32_t h = 0;
    HIPCHK(ctx, hipMemcpyAsync(&h, s->flags, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (h == 0) return MAVI_OK;
    HIPCHK(ctx, hipMemsetAsync(s->flags, 0, sizeof(uint32_t), ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (h & MAVI_FLAG_SCAN_STALL) return set_err(ctx, MAVI_E_HIP, "chunk-count scan: look-back did not complete");
    if (h & MAVI_FLAG_NEIGH_OVERFLOW)
        return set_err(ctx, MAVI_E_CELL_OVERFLOW,
                       "a particle has more than 15 contact neighbours (BoundsError in neigh_update_data!)");
    if (h & MAVI_FLAG_RING_OUT_OF_BOX)
        return set_err(ctx, MAVI_E_OUT_OF_BOX,
                       "a ring's centre of mass is outside the ring chunk grid (BoundsError in update_chunks! of "
                       "r_chunks)");
    if (h & MAVI_FLAG_RING_CHUNK_OVERFLOW)
        return set_err(ctx, MAVI_E_CELL_OVERFLOW,
                       "a ring chunk holds more rings than its capacity (BoundsError in update_chunks! of r_chunks)");
    if (h & MAVI_FLAG_NONFINITE)
        return set_err(ctx, MAVI_E_NONFINITE, "non-finite particle position met while binning");
    if (h & MAVI_FLAG_OUT_OF_BOX)
        return set_err(ctx, MAVI_E_OUT_OF_BOX, "particle outside the chunk grid (BoundsError in update_chunks!)");
    return set_err(ctx, MAVI_E_CELL_OVERFLOW,
                   "a chunk holds more particles than its capacity nc (BoundsError in update_chunks!)");
}

}  // namespace

double exact_gt_threshold(double rc) {
    if (!(rc > 0.0) || isinf(rc)) return INFINITY;
    double t = rc * rc;
    while (sqrt(t) > rc) t = nextafter(t, 0.0);
    while (sqrt(nextafter(t, INFINITY)) <= rc) t = nextafter(t, INFINITY);
    return t;
}

double exact_lt_threshold(double deq) {
    if (!(deq > 0.0)) return 0.0;
    double t = deq * deq;
    while (sqrt(t) < deq) t = nextafter(t, INFINITY);
    while (sqrt(nextafter(t, 0.0)) >= deq) t = nextafter(t, 0.0);
    return t;
}

extern "C" {

int mavi_abi_version(void) { return MAVI_ABI_VERSION; }

int mavi_ctx_create(int device, mavi_ctx **out) {
    if (!out) return MAVI_E_BADARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MAVI_E_HIP;
    if (device < 0 || device >= ndev) return MAVI_E_BADARG;
    mavi_ctx *ctx = new mavi_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MAVI_E_HIP;
    }
    // the no-alloc entry point's staging array on this device (first context
    // of the device only)
    if (mavi::noalloc_scratch_setup(device) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return MAVI_E_HIP;
    }
    *out = ctx;
    return MAVI_OK;
}

static void ctx_free(mavi_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void ctx_retain(mavi_ctx *ctx) { ctx->refs++; }

void ctx_release(mavi_ctx *ctx) {
    if (--ctx->refs == 0 && ctx->closing) ctx_free(ctx);
}

int mavi_ctx_destroy(mavi_ctx *ctx) {
    if (!ctx) return MAVI_OK;
    if (ctx->refs > 0) {  // systems, batches or networks still use its stream
        ctx->closing = true;
        return MAVI_OK;
    }
    ctx_free(ctx);
    return MAVI_OK;
}

const char *mavi_last_error(const mavi_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static void sys_free(mavi_sys *s);

void sys_retain(mavi_sys *s) { s->refs++; }

void sys_release(mavi_sys *s) {
    if (--s->refs == 0 && s->closing) sys_free(s);
}

int mavi_sys_destroy(mavi_sys *s) {
    if (!s) return MAVI_OK;
    if (s->refs > 0) {  // a learned propagator still steps it
        s->closing = true;
        return MAVI_OK;
    }
    sys_free(s);
    return MAVI_OK;
}

static void sys_free(mavi_sys *s) {
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    mavi::nl_free(s);
    if (s->graph2) (void)hipGraphExecDestroy(s->graph2);
    for (int c = 0; c < 2; c++) {
        dfree(s->P_[c]);
        dfree(s->V_[c]);
        dfree(s->ID_[c]);
    }
    dfree(s->F); dfree(s->P2); dfree(s->F1);
    dfree(s->key); dfree(s->rank); dfree(s->src); dfree(s->scell); dfree(s->sid); dfree(s->slot_of);
    dfree(s->cnt); dfree(s->start); dfree(s->scan_state); dfree(s->scan_ticket); dfree(s->flags); dfree(s->counter);
    dfree(s->stamp); dfree(s->epoch); dfree(s->perm); dfree(s->needc);
    dfree(s->active_ids); dfree(s->pos_id); dfree(s->vel_id); dfree(s->force_id);
    dfree(s->pol); dfree(s->areas); dfree(s->cms); dfree(s->cms_next); dfree(s->cpos);
    dfree(s->rtype); dfree(s->types); dfree(s->inter); dfree(s->noise); dfree(s->red);
    dfree(s->walls); dfree(s->lines); dfree(s->neigh_cnt); dfree(s->neigh_list);
    dfree(s->inv_cnt); dfree(s->inv_cell); dfree(s->inv_chunk); dfree(s->inv_box); dfree(s->inv_rec);
    dfree(s->inv_n);
    dfree(s->rmask); dfree(s->dsinks); dfree(s->dboxes); dfree(s->box_mask);
    if (s->vrb) (void)hipFree(s->vrb);
    if (s->vstage) (void)hipHostFree(s->vstage);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    mavi_ctx *ctx = s->ctx;
    const bool held = s->ctx_held;
    delete s;
    if (held) ctx_release(ctx);
}

}  // extern "C"

// Validate a descriptor and derive the kernel parameters and ring tables
// (System(...) / RingsSystem(...) construction, src/systems.jl:73-114,
// src/rings/rings.jl:231-291).  Shared by single systems and batches.
int build_params(const mavi_sys_desc *d, HostSys &h, std::string &err) {
    auto fail = [&](int code, const char *msg) {
        err = msg;
        return code;
    };
    if (d->n <= 0 || d->n > (int64_t)0x7fffffff / 2) return fail(MAVI_E_BADARG, "n must be in [1, 2^30)");
    if (!(d->size[0] > 0.0) || !(d->size[1] > 0.0)) return fail(MAVI_E_BADARG, "box size must be > 0");
    if (d->nwalls < 0 || (d->nwalls > 0 && !d->walls)) return fail(MAVI_E_BADARG, "walls list: nwalls < 0 or null");
    if (d->nwalls == 0 && d->wall != MAVI_WALL_RIGID && d->wall != MAVI_WALL_PERIODIC)
        return fail(MAVI_E_UNSUPPORTED, "unknown wall type");
    if (d->potential < MAVI_POT_LJ || d->potential > MAVI_POT_RTP)
        return fail(MAVI_E_UNSUPPORTED, "unknown potential kind");
    const bool active = d->potential == MAVI_POT_SZABO || d->potential == MAVI_POT_RTP;
    // walls list: the main (first) wall decides periodicity
    bool periodic = d->wall == MAVI_WALL_PERIODIC;
    bool rigid_listed = false;
    h.walls.clear();
    h.lines.clear();
    if (d->nwalls > 0) {
        if (d->nwalls > 64) return fail(MAVI_E_BADARG, "walls list longer than 64");
        periodic = d->walls[0].kind == MAVI_WK_PERIODIC_RECT;
        for (int k = 0; k < d->nwalls; k++) {
            const mavi_wall &w = d->walls[k];
            if (w.kind < MAVI_WK_RIGID_RECT || w.kind > MAVI_WK_POTENTIAL_LINES)
                return fail(MAVI_E_UNSUPPORTED, "unknown wall kind");
            if (w.kind == MAVI_WK_PERIODIC_RECT && k > 0)
                return fail(MAVI_E_UNSUPPORTED, "PeriodicWalls must be the main (first) wall");
            if ((w.kind == MAVI_WK_POTENTIAL_CIRCLE || w.kind == MAVI_WK_POTENTIAL_LINES) &&
                w.pot_kind != MAVI_POT_LJ && w.pot_kind != MAVI_POT_HARMTRUNC)
                return fail(MAVI_E_UNSUPPORTED, "potential walls take a LenJonesCfg or HarmTruncCfg");
            if ((w.kind == MAVI_WK_SLIPPERY_LINES || w.kind == MAVI_WK_POTENTIAL_LINES) && (w.nlines <= 0 || !w.lines))
                return fail(MAVI_E_BADARG, "LinesCfg wall without lines");
            rigid_listed = rigid_listed || w.kind == MAVI_WK_RIGID_RECT || w.kind == MAVI_WK_RIGID_CIRCLE;
            DevWall dw{};
            dw.kind = w.kind;
            dw.mode = w.mode;
            dw.pot_kind = w.pot_kind;
            for (int j = 0; j < 4; j++) {
                dw.g[j] = w.geom[j];
                dw.pot[j] = w.pot[j];
            }
            if (w.kind == MAVI_WK_SLIPPERY_LINES || w.kind == MAVI_WK_POTENTIAL_LINES) {
                dw.first = (int32_t)h.lines.size();
                dw.nlines = w.nlines;
                for (int l = 0; l < w.nlines; l++) {
                    // Line2D(p1, p2), src/configs.jl:99-117
                    const double *q = w.lines + 4 * l;
                    DevLine L{};
                    L.p1x = q[0]; L.p1y = q[1]; L.p2x = q[2]; L.p2y = q[3];
                    const double drx = q[2] - q[0], dry = q[3] - q[1];
                    const double norm = sqrt(drx * drx + dry * dry);
                    if (!(norm > 0.0)) return fail(MAVI_E_BADARG, "a line of zero length");
                    L.nx = -dry / norm; L.ny = drx / norm;
                    L.tx = drx / norm; L.ty = dry / norm;
                    L.len = norm;
                    h.lines.push_back(L);
                }
            }
            h.walls.push_back(dw);
        }
        if (periodic) {
            const double *g = d->walls[0].geom;
            if (g[0] != d->bottom_left[0] || g[1] != d->bottom_left[1] || g[2] != d->size[0] || g[3] != d->size[1])
                return fail(MAVI_E_UNSUPPORTED,
                            "a periodic main rectangle must be the bounding box of every geometry");
        }
    }
    if (active && (rigid_listed || (d->nwalls == 0 && !periodic)))
        return fail(MAVI_E_UNSUPPORTED,
                    "self-propelled states take no RigidWalls (walls! with RigidWalls reads state.vel, which "
                    "SelfPropelledState does not have)");
    if (d->potential == MAVI_POT_RINGS && rigid_listed)
        return fail(MAVI_E_UNSUPPORTED, "rings take no RigidWalls (walls! reads state.vel, which RingsState lacks)");
    if (d->potential == MAVI_POT_RINGS && d->nwalls > 0 && d->nmax > 32)
        return fail(MAVI_E_UNSUPPORTED, "rings with a walls list need nmax <= 32");
    const bool chunks = d->cols > 0 || d->rows > 0;
    if (chunks) {
        if (d->cols <= 0 || d->rows <= 0) return fail(MAVI_E_BADARG, "cols and rows must both be > 0");
        if (d->wall == MAVI_WALL_PERIODIC && (d->cols < 3 || d->rows < 3))
            return fail(MAVI_E_UNSUPPORTED,
                        "periodic chunk grids need >= 3 chunks per side (the reference stencil repeats "
                        "neighbours below that)");
        if ((int64_t)d->cols * d->rows > 0x3fffffff) return fail(MAVI_E_BADARG, "chunk grid too large");
        if (d->nc <= 0) return fail(MAVI_E_BADARG, "chunk capacity nc must be > 0");
    }
    if (!(d->dt > 0.0) && d->dt != 0.0) return fail(MAVI_E_BADARG, "dt must be >= 0");
    SysParams &P = h.P;
    P = SysParams{};
    P.n = (int32_t)d->n;
    P.cols = chunks ? d->cols : 0;
    P.rows = chunks ? d->rows : 0;
    P.ncell = P.cols * P.rows;
    P.nc = chunks ? d->nc : 0;
    P.periodic = periodic;
    P.nwalls = (int32_t)h.walls.size();
    P.potential = d->potential;
    P.blx = d->bottom_left[0];
    P.bly = d->bottom_left[1];
    P.L = d->size[0];
    P.H = d->size[1];
    P.cl = chunks ? d->size[0] / (double)P.cols : 0.0;  // src/chunks.jl:29-30
    P.ch = chunks ? d->size[1] / (double)P.rows : 0.0;
    P.halfL = d->size[0] / 2.0;
    P.halfH = d->size[1] / 2.0;
    P.dt = d->dt;
    P.term = d->dt * d->dt / 2.0;
    P.hdt = d->dt / 2.0;
    P.radius = d->radius;
    P.p0 = d->pot[0];
    P.p1 = d->pot[1];
    P.p2 = d->pot[2];
    P.p3 = d->pot[3];
    P.cut2 = INFINITY;
    P.eq2 = 0.0;
    h.list2 = INFINITY;
    if (d->potential == MAVI_POT_LJ) {
        P.cut2 = exact_gt_threshold(d->pot[2]);
        const double s2 = d->pot[0] * d->pot[0], s6 = s2 * s2 * s2;
        P.lj_c12 = 48.0 * d->pot[1] * s6 * s6;
        P.lj_c6 = 24.0 * d->pot[1] * s6;
        h.list2 = P.cut2;
    } else if (d->potential == MAVI_POT_HARMTRUNC) {
        P.cut2 = exact_gt_threshold(d->pot[3]);
        P.eq2 = exact_lt_threshold(d->pot[2]);
        h.list2 = P.cut2;
    } else if (d->potential == MAVI_POT_SZABO) {
        // pot: k_rep, k_adh, r_eq, r_max; active: vo, mobility, relax_time, rot_diff
        if (!(d->pot[3] > d->pot[2]) || !(d->pot[2] > 0.0) || !(d->active[2] > 0.0))
            return fail(MAVI_E_BADARG, "SzaboCfg needs 0 < r_eq < r_max and relax_time > 0");
        P.cut2 = exact_gt_threshold(d->pot[3]);
        P.sz_rep = d->pot[0] / (d->pot[3] - d->pot[2]);
        P.sz_adh = d->pot[1] / d->pot[2];
        P.ap_vo = d->active[0];
        P.ap_mu = d->active[1];
        P.ap_inv_tau = 1.0 / d->active[2];
        P.ap_noise = sqrt(2.0 * d->active[3] * d->dt);
        h.list2 = P.cut2;
    } else if (d->potential == MAVI_POT_RTP) {
        // pot: sigma, epsilon; the force is the LJ one cut at 2^(1/6) sigma
        // (src/integration.jl:88-109); active: vo, tumble_rate
        const double cutoff = pow(2.0, 1.0 / 6.0) * d->pot[0];
        P.p2 = cutoff;
        P.cut2 = exact_gt_threshold(cutoff);
        const double s2 = d->pot[0] * d->pot[0], s6 = s2 * s2 * s2;
        P.lj_c12 = 48.0 * d->pot[1] * s6 * s6;
        P.lj_c6 = 24.0 * d->pot[1] * s6;
        P.ap_vo = d->active[0];
        P.ap_tumble = d->active[1] * d->dt;
        h.list2 = P.cut2;
    }
    h.active.clear();
    h.rtype.clear();
    h.types.clear();
    h.inter.clear();
    h.seed = d->seed;
    if (d->potential == MAVI_POT_RINGS) {
        if (!periodic) return fail(MAVI_E_UNSUPPORTED, "rings run on the GPU with a periodic main wall");
        if (d->nmax <= 0 || d->nrings <= 0 || (int64_t)d->nmax * d->nrings != d->n || d->ntypes <= 0 ||
            !d->type_np || !d->type_par || !d->interaction)
            return fail(MAVI_E_BADARG, "rings descriptor: need n == nmax*nrings and type tables");
        P.nmax = d->nmax;
        P.nrings = d->nrings;
        P.ntypes = d->ntypes;
        P.typed = d->ring_types != nullptr;
        if (!P.typed && d->ntypes != 1) return fail(MAVI_E_BADARG, "untyped rings need ntypes == 1");
        for (int t = 0; t < d->ntypes; t++) {
            RingType T{};
            const double *q = d->type_par + 8 * t;
            T.p0 = q[0]; T.relax = q[1]; T.vo = q[2]; T.mu = q[3];
            T.rot = q[4]; T.karea = q[5]; T.kspring = q[6]; T.lspring = q[7];
            T.np = d->type_np[t];
            if (T.np < 1 || T.np > d->nmax) return fail(MAVI_E_BADARG, "ring particle count outside [1, nmax]");
            // get_particle_radius: the type's own interaction, dist_eq / 2
            T.wradius = d->interaction[4 * (t + d->ntypes * t) + 2] / 2.0;
            h.types.push_back(T);
        }
        h.list2 = 0.0;
        for (int k = 0; k < d->ntypes * d->ntypes; k++) {
            RingInter I{};
            const double *q = d->interaction + 4 * k;
            I.k_rep = q[0]; I.k_atr = q[1]; I.dist_eq = q[2]; I.dist_max = q[3];
            I.max2 = exact_gt_threshold(q[3]);
            I.eq2 = exact_lt_threshold(q[2]);
            if (I.max2 > h.list2) h.list2 = I.max2;
            h.inter.push_back(I);
        }
        h.rtype.resize(d->nrings);
        for (int r = 0; r < d->nrings; r++) {
            int t = P.typed ? d->ring_types[r] - 1 : 0;
            if (t < 0 || t >= d->ntypes) return fail(MAVI_E_BADARG, "ring type outside [1, ntypes]");
            h.rtype[r] = t;
            // FixRingsIds.p_ids (src/rings/states.jl:46-62): the first
            // num_particles[type] slots of every ring, ascending.
            for (int i = 0; i < h.types[t].np; i++) h.active.push_back(r * d->nmax + i);
        }
        P.na = (int32_t)h.active.size();
    } else {
        P.na = P.n;
    }
    return MAVI_OK;
}

extern "C" {

int mavi_sys_create(mavi_ctx *ctx, const mavi_sys_desc *d, mavi_sys **out) {
    if (!ctx || !d || !out) return MAVI_E_BADARG;
    *out = nullptr;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HostSys h;
    std::string err;
    int st = build_params(d, h, err);
    if (st != MAVI_OK) return set_err(ctx, st, err);
    mavi_sys *s = new mavi_sys();
    s->ctx = ctx;
    s->P = h.P;
    s->list2 = h.list2;
    SysParams &P = s->P;
    std::vector<int32_t> &active = h.active;
    std::vector<int32_t> &host_rtype = h.rtype;
    std::vector<RingType> &host_types = h.types;
    std::vector<RingInter> &host_inter = h.inter;

    hipError_t e = hipSuccess;
#define ALLOC(ptr, cnt) \
    if (e == hipSuccess) e = dalloc(&(ptr), (cnt), ctx->stream)
    const bool rings = P.potential == MAVI_POT_RINGS;
    for (int c = 0; c < 2; c++) {
        ALLOC(s->P_[c], P.na);
        if (!rings) ALLOC(s->V_[c], P.na);
        ALLOC(s->ID_[c], P.na);
    }
    ALLOC(s->F, P.na);
    ALLOC(s->P2, P.na);
    ALLOC(s->F1, P.na);
    ALLOC(s->key, P.na);
    ALLOC(s->rank, P.na);
    ALLOC(s->src, P.na);
    ALLOC(s->scell, P.na);
    ALLOC(s->sid, P.na);
    ALLOC(s->slot_of, P.n);
    ALLOC(s->flags, 1);
    ALLOC(s->stamp, (int64_t)P.ncell + 1);
    ALLOC(s->perm, P.na);
    ALLOC(s->needc, (int64_t)P.ncell + 1);
    ALLOC(s->epoch, 1);
    ALLOC(s->counter, 1);
    ALLOC(s->cnt, (int64_t)P.ncell + 1);
    ALLOC(s->start, (int64_t)P.ncell + 1);
    // one state word per scan tile (tiles hold at least 4096 counts)
    s->nscan_blocks = (P.ncell + 1 + 4095) / 4096;
    ALLOC(s->scan_state, s->nscan_blocks);
    ALLOC(s->scan_ticket, 1);
    ALLOC(s->pos_id, P.n);
    if (!rings) ALLOC(s->vel_id, P.n);
    ALLOC(s->force_id, P.n);
    if (rings) {
        ALLOC(s->active_ids, P.na);
        ALLOC(s->pol, P.nrings);
        ALLOC(s->areas, P.nrings);
        ALLOC(s->cms, P.nrings);
        ALLOC(s->cms_next, P.nrings);
        ALLOC(s->cpos, P.n);
        ALLOC(s->rtype, P.nrings);
        ALLOC(s->types, P.ntypes);
        ALLOC(s->inter, (int64_t)P.ntypes * P.ntypes);
    }
    if (!h.walls.empty()) {
        ALLOC(s->walls, (int64_t)h.walls.size());
        if (!h.lines.empty()) ALLOC(s->lines, (int64_t)h.lines.size());
    }
#undef ALLOC
    if (e == hipSuccess && !h.walls.empty()) {
        e = hipMemcpyAsync(s->walls, h.walls.data(), sizeof(DevWall) * h.walls.size(), hipMemcpyHostToDevice,
                           ctx->stream);
        if (e == hipSuccess && !h.lines.empty())
            e = hipMemcpyAsync(s->lines, h.lines.data(), sizeof(DevLine) * h.lines.size(), hipMemcpyHostToDevice,
                               ctx->stream);
        P.walls = s->walls;
        P.lines = s->lines;
    }
    if (e == hipSuccess) e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e == hipSuccess && P.potential == MAVI_POT_RINGS) {
        auto up = [&](void *dst, const void *src, size_t bytes) {
            return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
        };
        e = up(s->active_ids, active.data(), sizeof(int32_t) * active.size());
        if (e == hipSuccess) e = up(s->rtype, host_rtype.data(), sizeof(int32_t) * host_rtype.size());
        if (e == hipSuccess) e = up(s->types, host_types.data(), sizeof(RingType) * host_types.size());
        if (e == hipSuccess) e = up(s->inter, host_inter.data(), sizeof(RingInter) * host_inter.size());
        s->host_inter = host_inter;
        s->host_ring_np.resize(P.nrings);
        for (int32_t r = 0; r < P.nrings; r++) s->host_ring_np[r] = host_types[P.typed ? host_rtype[r] : 0].np;
    }
    // neighbour lists for finite-range plain systems (at most 2^31 list entries)
    if (e == hipSuccess && mavi::nl_eligible(s) && (int64_t)s->P.na * 64 < (int64_t)INT32_MAX)
        e = mavi::nl_alloc(s, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    s->seed = h.seed;
    if (e != hipSuccess) {
        std::string m = std::string("allocation/upload failed: ") + hipGetErrorString(e);
        mavi_sys_destroy(s);
        return set_err(ctx, MAVI_E_HIP, m);
    }
    ctx_retain(ctx);
    s->ctx_held = true;
    *out = s;
    return MAVI_OK;
}

int mavi_sys_upload(mavi_sys *s, const double *pos, const double *vel, const double *pol) {
    if (!s || !pos) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double2) * (size_t)s->P.n;
    HIPCHK(ctx, hipMemcpyAsync(s->pos_id, pos, bytes, hipMemcpyHostToDevice, ctx->stream));
    if (is_active(s)) {
        // pol_angle (n) -> the x components of the velocity staging
        HIPCHK(ctx, hipMemsetAsync(s->vel_id, 0, bytes, ctx->stream));
        if (pol)
            HIPCHK(ctx, hipMemcpy2DAsync(s->vel_id, sizeof(double2), pol, sizeof(double), sizeof(double),
                                         (size_t)s->P.n, hipMemcpyHostToDevice, ctx->stream));
    } else if (vel && s->vel_id) {
        HIPCHK(ctx, hipMemcpyAsync(s->vel_id, vel, bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    if (pol && s->P.potential == MAVI_POT_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(s->pol, pol, sizeof(double) * s->P.nrings, hipMemcpyHostToDevice, ctx->stream));
    s->cur = 0;
    if (s->nl) s->nl->valid = false;  // a new state in id order
    HIPCHK(ctx, mavi::drop_keys(s, ctx->stream));
    HIPCHK(ctx, mavi::launch_iota_state(s, ctx->stream));
    if (s->P.potential == MAVI_POT_RINGS) HIPCHK(ctx, mavi::launch_rings_init(s, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    s->has_state = true;
    s->f_valid = false;
    s->sorted = false;
    return MAVI_OK;
}

// Bring the id-order staging copies up to date with the sorted state.
static int sync_staging(mavi_sys *s, bool forces) {
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, mavi::launch_unpermute(s, true, forces, ctx->stream));
    return MAVI_OK;
}

int mavi_sys_download(mavi_sys *s, double *pos, double *vel, double *pol) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double2) * (size_t)s->P.n;
    if (pos) HIPCHK(ctx, hipMemcpyAsync(pos, s->pos_id, bytes, hipMemcpyDeviceToHost, ctx->stream));
    if (is_active(s)) {
        if (pol)
            HIPCHK(ctx, hipMemcpy2DAsync(pol, sizeof(double), s->vel_id, sizeof(double2), sizeof(double),
                                         (size_t)s->P.n, hipMemcpyDeviceToHost, ctx->stream));
    } else if (vel && s->vel_id) {
        HIPCHK(ctx, hipMemcpyAsync(vel, s->vel_id, bytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (pol && s->P.potential == MAVI_POT_RINGS)
        HIPCHK(ctx, hipMemcpyAsync(pol, s->pol, sizeof(double) * s->P.nrings, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MAVI_OK;
}

int mavi_bin(mavi_sys *s, int64_t *chunk_particles, int64_t *counts) {
    if (s) s->f_valid = false;  // a binning the next step's stamps would not see
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (s->P.cols == 0) return set_err(ctx, MAVI_E_BADARG, "system has no chunks");
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "no state uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, mavi::launch_bin(s, ctx->stream));
    int st = sync_staging(s, false);  // binning moves no force
    if (st == MAVI_OK) st = check_flags(s);
    if (st != MAVI_OK) return st;
    if (!chunk_particles && !counts) return MAVI_OK;
    const int64_t ncell = s->P.ncell, nc = s->P.nc;
    int64_t *dcp = nullptr, *dcnt = nullptr;
    HIPCHK(ctx, dalloc(&dcp, nc * ncell, ctx->stream));
    hipError_t e = dalloc(&dcnt, ncell, ctx->stream);
    if (e != hipSuccess) { dfree(dcp); HIPCHK(ctx, e); }
    if (chunk_particles)
        e = hipMemcpyAsync(dcp, chunk_particles, sizeof(int64_t) * nc * ncell, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = mavi::launch_export_chunks(s, dcp, dcnt, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && chunk_particles)
        e = hipMemcpy(chunk_particles, dcp, sizeof(int64_t) * nc * ncell, hipMemcpyDeviceToHost);
    if (e == hipSuccess && counts) e = hipMemcpy(counts, dcnt, sizeof(int64_t) * ncell, hipMemcpyDeviceToHost);
    dfree(dcp);
    dfree(dcnt);
    HIPCHK(ctx, e);
    return MAVI_OK;
}

int mavi_forces(mavi_sys *s, double *forces) {
    if (s) s->f_valid = false;
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "no state uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (s->P.potential == MAVI_POT_RINGS) {
        HIPCHK(ctx, mavi::launch_rings_forces(s, ctx->stream));
    } else {
        HIPCHK(ctx, mavi::launch_bin(s, ctx->stream));
        HIPCHK(ctx, mavi::launch_forces_only(s, ctx->stream));
    }
    int st = sync_staging(s, true);
    if (st == MAVI_OK) st = check_flags(s);
    if (st != MAVI_OK) return st;
    if (forces)
        HIPCHK(ctx, hipMemcpy(forces, s->force_id, sizeof(double2) * (size_t)s->P.n, hipMemcpyDeviceToHost));
    return MAVI_OK;
}

int mavi_get_forces(mavi_sys *s, double *forces) {
    if (!s || !forces) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(forces, s->force_id, sizeof(double2) * (size_t)s->P.n, hipMemcpyDeviceToHost));
    return MAVI_OK;
}

// One step of the system's dynamics (newton_step! or Rings step!) with the
// device rotational noise.  reuse: the first force evaluation may take the
// previous step's second one where it is provably identical.
static hipError_t launch_step(mavi_sys *s, hipStream_t st, bool reuse) {
    if (s->P.potential == MAVI_POT_RINGS) return mavi::launch_rings_step(s, nullptr, st);
    if (is_active(s)) return mavi::launch_active_step(s, nullptr, st);
    return mavi::launch_verlet_step(s, st, reuse);
}

// Two steps captured once into a graph (state starting in copy 0, so the
// ping-pong returns to copy 0) and replayed: one launch per two steps.  The
// graph's steps reuse forces, so it is replayed only after a Verlet step.
static hipError_t ensure_graph(mavi_sys *s, hipStream_t st) {
    if (s->graph2) return hipSuccess;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const int cur0 = s->cur;
    hipError_t e1 = launch_step(s, st, true);
    hipError_t e2 = (e1 == hipSuccess) ? launch_step(s, st, true) : e1;
    e = hipStreamEndCapture(st, &g);
    s->cur = cur0;
    if (e1 != hipSuccess) e = e1;
    if (e == hipSuccess && e2 != hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&s->graph2, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

// check_invasions! on the device, then the record count read back: the
// record buffer grows and the tests rerun when a check finds more records
// than it holds.  Ring-grid errors (BoundsError in the reference) end the
// call here.
static int run_invasions(mavi_sys *s) {
    mavi_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, mavi::launch_invasions(s, true, st));
    for (;;) {
        uint32_t n = 0;
        HIPCHK(ctx, hipMemcpyAsync(&n, s->inv_n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        const int rc = check_flags(s);
        if (rc != MAVI_OK) {
            s->inv_count = 0;
            return rc;
        }
        if ((int64_t)n <= s->inv_cap) {
            s->inv_count = n;
            return MAVI_OK;
        }
        dfree(s->inv_rec);
        s->inv_cap = 0;
        HIPCHK(ctx, dalloc(&s->inv_rec, 3 * (int64_t)n, st));
        s->inv_cap = n;
        HIPCHK(ctx, mavi::launch_invasions(s, false, st));
    }
}

// ---------------------------------------------------------------------------
// Variable ring populations: update_sources! (src/rings/integration.jl:
// 353-358, src/rings/sources.jl:187-260) and update_ids! at the start of a
// step.  The device computes update_cms!, the sinks holding each ring
// slot's centre of mass and the occupied spawn boxes; the host replays the
// reference's sequential logic on its copy of VarRingsIds (list order, the
// lowest free slot, the stale tail of ids); a changed population is pushed
// back and the sorted state rebuilt from the id-order staging.
// ---------------------------------------------------------------------------
constexpr uint32_t kPolStreamId = 0xFFFFFFFFu;  // the spawn polarisation stream's id word

struct VarReadback {  // the layout of the device read-back buffer's first bytes
    uint32_t sink_n, pad;
    unsigned long long nonempty;
    int32_t list[2 * kSinkFetch];
};
constexpr size_t kVarHeader = 16;
static_assert(offsetof(VarReadback, nonempty) == 8 && offsetof(VarReadback, list) == kVarHeader,
              "read-back header layout");

static size_t var_stage_bytes(const mavi_sys *s) {
    const size_t spawn = (size_t)kMaxBoxes * (2 * (size_t)s->P.nmax + 3) * sizeof(double);
    return sizeof(VarReadback) + sizeof(int32_t) * ((size_t)s->P.nrings + (size_t)s->P.n) + spawn + 256;
}

// calc_active_ids! (src/rings/states.jl:200-223) on the host copy; active
// receives the p_ids (0-based slots)
static void var_update_ids(mavi_sys *s, std::vector<int32_t> &active) {
    int64_t k = 0;
    active.clear();
    for (int32_t r = 0; r < s->P.nrings; r++) {
        if (!s->vmask[r]) continue;
        s->vids[k++] = r + 1;
        for (int32_t i = 0; i < s->host_ring_np[r]; i++) active.push_back(r * s->P.nmax + i);
    }
    s->vnum_active = k;
}

// Push the host mask and active ids, the spawned rings and rebuild the state.
struct VarSpawn {
    int32_t slot, src, box;
    double pol, cmx, cmy;
};
static int var_apply(mavi_sys *s, const std::vector<VarSpawn> &spawns, const std::vector<int32_t> &active) {
    mavi_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    char *base = (char *)s->vstage + sizeof(VarReadback);
    int32_t *hmask = (int32_t *)base;
    int32_t *hact = hmask + s->P.nrings;
    double *hsp = (double *)(((uintptr_t)(hact + s->P.n) + 15) & ~(uintptr_t)15);
    // positions of the current active particles back to id order
    HIPCHK(ctx, mavi::launch_unpermute(s, false, false, st));
    const int64_t nm = s->P.nmax;
    for (size_t k = 0; k < spawns.size(); k++) {
        const VarSpawn &v = spawns[k];
        double *d = hsp + k * (2 * nm + 3);
        memcpy(d, s->vsrc[v.src].spawn.data() + (size_t)v.box * 2 * nm, sizeof(double) * 2 * nm);
        d[2 * nm] = v.pol;
        d[2 * nm + 1] = v.cmx;
        d[2 * nm + 2] = v.cmy;
        HIPCHK(ctx, hipMemcpyAsync(s->pos_id + (int64_t)v.slot * nm, d, sizeof(double) * 2 * nm,
                                   hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(s->pol + v.slot, d + 2 * nm, sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(s->cms + v.slot, d + 2 * nm + 1, sizeof(double2), hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipMemcpyAsync(s->cms_next + v.slot, d + 2 * nm + 1, sizeof(double2), hipMemcpyHostToDevice, st));
    }
    memcpy(hmask, s->vmask.data(), sizeof(int32_t) * s->P.nrings);
    HIPCHK(ctx, hipMemcpyAsync(s->rmask, hmask, sizeof(int32_t) * s->P.nrings, hipMemcpyHostToDevice, st));
    if (!active.empty()) {
        memcpy(hact, active.data(), sizeof(int32_t) * active.size());
        HIPCHK(ctx, hipMemcpyAsync(s->active_ids, hact, sizeof(int32_t) * active.size(), hipMemcpyHostToDevice, st));
    }
    s->P.na = (int32_t)active.size();
    HIPCHK(ctx, mavi::launch_iota_state(s, st));
    s->sorted = false;
    return MAVI_OK;
}

static bool host_inside_sink(const HostSource &q, double x, double y) {
    if (q.geom == MAVI_GEOM_RECT)
        return (q.g[0] - 0.0 <= x && x <= q.g[0] + q.g[2] + 0.0) && (q.g[1] - 0.0 <= y && y <= q.g[1] + q.g[3] + 0.0);
    const double dx = x - q.g[0], dy = y - q.g[1];
    return dx * dx + dy * dy <= q.g[2] * q.g[2];
}

static int var_step_start(mavi_sys *s) {
    mavi_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, mavi::drop_keys(s, st));  // keys of steps before the population became variable
    VarReadback *rb = (VarReadback *)s->vstage;
    HIPCHK(ctx, mavi::launch_var_pre(s, st));
    const int64_t fetch = s->P.nrings < kSinkFetch ? s->P.nrings : kSinkFetch;  // the list holds 2 x nrings
    HIPCHK(ctx, hipMemcpyAsync(rb, s->vrb, kVarHeader + sizeof(int32_t) * 2 * fetch, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    const uint32_t nhit = rb->sink_n;
    std::vector<uint32_t> &bits = s->vbits;
    std::vector<int32_t> extra;
    if ((int64_t)nhit > fetch) {
        extra.resize(2 * (size_t)nhit);
        HIPCHK(ctx, hipMemcpy(extra.data(), s->sink_list, sizeof(int32_t) * 2 * nhit, hipMemcpyDeviceToHost));
    }
    const int32_t *lst = extra.empty() ? rb->list : extra.data();
    for (uint32_t k = 0; k < nhit; k++) bits[lst[2 * k]] = (uint32_t)lst[2 * k + 1];
    const unsigned long long nonempty = rb->nonempty;
    // the reference's sequential logic, in list order
    std::vector<VarSpawn> spawns;
    std::vector<char> &added = s->vadded;
    std::vector<double> &acm = s->vacm;
    bool changed = false;
    for (size_t q = 0; q < s->vsrc.size(); q++) {
        const HostSource &src = s->vsrc[q];
        if (src.kind == MAVI_SRC_SINK) {
            const int64_t nview = s->vnum_active;  // get_rings_ids(state) when the loop starts
            for (int64_t k = 0; k < nview; k++) {
                const int64_t r = s->vids[k] - 1;
                const bool in = added[r] ? host_inside_sink(src, acm[2 * r], acm[2 * r + 1])
                                         : ((bits[r] >> src.sink_bit) & 1u) != 0;
                if (in) {  // remove_ring!
                    s->vmask[r] = 0;
                    s->vnum_active -= 1;
                    changed = true;
                }
            }
            continue;
        }
        const int64_t nm = s->P.nmax;
        for (int32_t b = 0; b < src.nbox; b++) {
            if ((nonempty >> (src.box0 + b)) & 1ull) continue;
            double pol = src.pol;  // get_spawn_pol runs before add_ring! looks for a slot
            if (src.random_pol) {
                double u, w;
                if (s->pol_used < (int64_t)s->pol_draws.size()) {
                    u = s->pol_draws[s->pol_used++];
                } else {  // the built-in stream: Philox4x32-10 keyed by the system's seed
                    philox_uniform2(s->seed, (uint64_t)s->pol_stream++, kPolStreamId, &u, &w);
                }
                pol = u * 2.0 * 3.141592653589793;  // rand(rng, T) * 2 * pi
            }
            int32_t slot = -1;
            for (int32_t r = 0; r < s->P.nrings; r++)
                if (!s->vmask[r]) { slot = r; break; }
            if (slot < 0) continue;  // add_ring! found no room
            s->vmask[slot] = 1;
            int64_t umax = s->vuids[0];
            for (int32_t r = 1; r < s->P.nrings; r++) umax = s->vuids[r] > umax ? s->vuids[r] : umax;
            s->vuids[slot] = umax + 1;
            s->vnum_active += 1;
            // info.cms[ring_id] = sum(rings_pos[1:num_p, ring_id]) / num_p
            const double *p = src.spawn.data() + (size_t)b * 2 * nm;
            const int32_t np = s->host_ring_np[slot];
            double sx = 0.0, sy = 0.0;
            for (int32_t i = 0; i < np; i++) {
                sx += p[2 * i];
                sy += p[2 * i + 1];
            }
            acm[2 * slot] = sx / (double)np;
            acm[2 * slot + 1] = sy / (double)np;
            added[slot] = 1;
            spawns.push_back(VarSpawn{slot, (int32_t)q, b, pol, acm[2 * slot], acm[2 * slot + 1]});
            changed = true;
        }
    }
    for (uint32_t k = 0; k < nhit; k++) bits[lst[2 * k]] = 0u;  // scratch back to zero
    for (const VarSpawn &v : spawns) added[v.slot] = 0;
    // an unchanged mask leaves update_ids!'s result as it is
    if (!changed) return MAVI_OK;
    std::vector<int32_t> active;
    var_update_ids(s, active);
    return var_apply(s, spawns, active);
}

// update_invasions! (src/rings/integration.jl:511-520) at the start of a
// step: num_steps - last_check >= steps_to_update.
static bool invasions_due(const mavi_sys *s, int64_t ahead) {
    return s->inv_sched && s->inv_since + ahead >= s->inv_every;
}

// Neighbour-list stepping of a finite-range plain system: blocks of steps
// (six-step graphs), after each block one read back of the largest drift
// step and of the largest displacement of the next positions since the last
// build.  A block in which some particle may have moved more than skin / 2
// may have missed a pair: its start state is restored, the lists are rebuilt
// there and the block runs again (shorter).  A block runs as many steps as
// the bound allows at the last block's step length (block option 0), or a
// fixed count; the lists are rebuilt ahead when fewer than kNlMinBlock more
// steps (fixed: the next block at 1.2 times the step length) would fit.
static bool nl_active(const mavi_sys *s) { return s->nl && s->nl->on && !s->nl->disabled; }

static hipError_t nl_graph(mavi_sys *s, hipStream_t st) {
    NList &N = *s->nl;
    hipGraphExec_t &gx = N.graph[s->cur];  // the kernels hold the state copy's buffers
    if (gx) return hipSuccess;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const int ph0 = N.ph;
    hipError_t e1 = hipSuccess;
    for (int k = 0; k < 6 && e1 == hipSuccess; k++) e1 = mavi::nl_step(s, st);
    e = hipStreamEndCapture(st, &g);
    N.ph = ph0;
    if (e1 != hipSuccess) e = e1;
    if (e == hipSuccess) e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

static void nl_drop_graph(mavi_sys *s) {
    if (!s->nl) return;
    for (auto &g : s->nl->graph) {
        if (!g) continue;
        (void)hipStreamSynchronize(s->ctx->stream);
        (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

constexpr int64_t kNlMinBlock = 4;  // shorter blocks cost more than the rebuild they put off

// The largest drift step of the last three blocks (0: none measured yet).
static double nl_step_est(const NList &N) {
    return std::max(N.step_hist[0], std::max(N.step_hist[1], N.step_hist[2]));
}

// Steps from a block start (d_start from the build, `fresh`: the block starts
// with the build's drift) that keep d_start + (steps - 1 + fresh) x step
// within half_skin, at a step 1.08 times the estimate.
static int64_t nl_room(const NList &N, double half_skin, bool fresh) {
    const double r = (half_skin - N.d_start) / (1.08 * nl_step_est(N));
    if (!(r < 1e9)) return INT64_MAX;
    return (int64_t)floor(r) + (fresh ? 0 : 1);
}

static int nl_steps(mavi_sys *s, int64_t nsteps, int64_t *left) {
    mavi_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    NList &N = *s->nl;
    const SysParams &P = s->P;
    const size_t bytes = sizeof(double2) * (size_t)P.na;
    const double half_skin = sqrt(N.thr2);
    *left = 0;
    int64_t k = 0;
    while (k < nsteps) {
        const bool fresh = !N.valid;
        if (fresh) HIPCHK(ctx, mavi::nl_rebuild(s, st));  // resets the step maxima, then the first drift
        int64_t want = N.block_now;
        if (N.block == 0) {
            if (nl_step_est(N) > 0.0) want = std::min(want, std::max<int64_t>(1, nl_room(N, half_skin, fresh)));
            else want = std::min<int64_t>(want, 8);  // no step length measured yet
        }
        const int64_t blk = std::min<int64_t>(want, nsteps - k);
        // the block's start state, to roll back to (the state's positions are
        // written only after the check)
        HIPCHK(ctx, mavi::nl_block_start(s, st, !fresh));
        HIPCHK(ctx, hipMemcpyAsync(N.vsnap, s->V_[s->cur], bytes, hipMemcpyDeviceToDevice, st));
        int64_t done = 0;
        const bool graphs = s->use_graphs && !s->prof;
        while (done < blk) {
            if (graphs && N.ph == 0 && done + 6 <= blk) {
                HIPCHK(ctx, nl_graph(s, st));
                HIPCHK(ctx, hipGraphLaunch(N.graph[s->cur], st));
                N.unsettled = true;
                done += 6;
            } else {
                HIPCHK(ctx, mavi::nl_step(s, st));
                done += 1;
            }
        }
        HIPCHK(ctx, mavi::nl_measure(s, st));
        HIPCHK(ctx, hipMemcpyAsync(N.host, N.dmax, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        const double step = sqrt(__builtin_bit_cast(double, N.host[0]));
        const double d_end = sqrt(__builtin_bit_cast(double, N.host[1]));
        const uint32_t status = (uint32_t)N.host[2];
        N.step_hist[N.step_i] = step;
        N.step_i = (N.step_i + 1) % 3;
        // The block evaluated positions y_0 .. y_{T-1} and drifted to y_T, each
        // drift at most `step` long; y_0 is within d_start of the build
        // positions (a block after a build starts with the build positions
        // y_{-1}, d_start = 0), y_T within d_end.  So y_t is within
        // min(d_start + u step, d_end + (U - u) step), u = t (+1 after a
        // build), U = T (+1): at most the two lines' crossing, and at most
        // d_start + (U - 1) step.
        const double U = (double)blk + (fresh ? 1.0 : 0.0);
        const double bound = std::min(N.d_start + (U - 1.0) * step, 0.5 * (N.d_start + d_end + U * step));
        if (status || bound > half_skin) {
            // a step of this block may have used an incomplete list: back to its start
            HIPCHK(ctx, hipMemcpyAsync(s->V_[s->cur], N.vsnap, bytes, hipMemcpyDeviceToDevice, st));
            HIPCHK(ctx, hipMemcpyAsync(s->flags, N.fsnap, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            N.valid = false;
            N.unsettled = false;  // the state's positions are the block start's
            N.rollbacks++;
            // a list overflowed, a chunk count may have passed the capacity
            // (the chunk gather checks it exactly), or one step from a fresh
            // build moved a particle more than skin / 2
            if (status || (fresh && blk == 1)) {
                // the chunk gather takes over
                N.disabled = true;
                HIPCHK(ctx, hipMemsetAsync(N.status, 0, sizeof(uint32_t), st));
                HIPCHK(ctx, hipStreamSynchronize(st));
                s->sorted = false;
                s->f_valid = false;
                *left = nsteps - k;  // for the caller's chunk gather
                return MAVI_OK;
            }
            // (a block sized by the step estimate is shorter next time: the
            // estimate now holds this block's step)
            if (N.block != 0 && N.block_now > 1) N.block_now = std::max(1, N.block_now / 2);
            continue;
        }
        HIPCHK(ctx, mavi::nl_settle(s, st));
        k += blk;
        N.d_start = d_end;
        // rebuild ahead when too few steps would fit before the bound
        if (N.block == 0) {
            if (nl_room(N, half_skin, false) < kNlMinBlock) N.valid = false;
        } else if (d_end + 1.2 * step * (double)N.block_now > half_skin) {
            N.valid = false;
        }
        const int32_t cap = N.block == 0 ? kNlBlockCap : N.block;
        if (N.block_now < cap) N.block_now = std::min(cap, 2 * N.block_now);
    }
    s->sorted = false;   // not in reference chunk order
    s->f_valid = false;  // the chunk gather's reuse does not apply to this order
    return MAVI_OK;
}

static int steps_impl(mavi_sys *s, int64_t nsteps, const double *noise_host) {
    mavi_ctx *ctx = s->ctx;
    if (nsteps < 0) return set_err(ctx, MAVI_E_BADARG, "nsteps < 0");
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "no state uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // a due invasion check runs after the step's rings kernel, which leaves
    // cms and cpos as the reference's step sees them before its update
    auto before_step = [&]() -> int { return s->var ? var_step_start(s) : MAVI_OK; };
    auto after_step = [&](bool due) -> int {
        if (!s->inv_sched) return MAVI_OK;
        s->inv_since = due ? 1 : s->inv_since + 1;
        return due ? run_invasions(s) : MAVI_OK;
    };
    if (!noise_host && nl_active(s) && nsteps > 0) {
        int64_t left = 0;
        const int rc = nl_steps(s, nsteps, &left);
        if (rc != MAVI_OK) return rc;
        if (left == 0) {
            int rc2 = sync_staging(s, true);
            if (rc2 != MAVI_OK) return rc2;
            return check_flags(s);
        }
        nsteps = left;  // lists disabled (a list overflowed): the chunk gather runs the rest
    }
    if (noise_host) {
        // draws per step: one per ring (rings), one (Szabo) or two (RTP) per particle
        const int64_t nr = s->P.potential == MAVI_POT_RINGS ? s->P.nrings
                           : s->P.potential == MAVI_POT_RTP ? 2 * (int64_t)s->P.n
                                                             : s->P.n;
        const int64_t need = nsteps * nr;
        if (need > s->noise_cap) {
            dfree(s->noise);
            s->noise_cap = 0;
            HIPCHK(ctx, dalloc(&s->noise, need, st));
            s->noise_cap = need;
        }
        HIPCHK(ctx, hipMemcpyAsync(s->noise, noise_host, sizeof(double) * need, hipMemcpyHostToDevice, st));
        for (int64_t k = 0; k < nsteps; k++) {
            const bool due = invasions_due(s, 0);
            const int rc0 = before_step();
            if (rc0 != MAVI_OK) return rc0;
            if (s->P.potential == MAVI_POT_RINGS) HIPCHK(ctx, mavi::launch_rings_step(s, s->noise + k * nr, st));
            else HIPCHK(ctx, mavi::launch_active_step(s, s->noise + k * nr, st));
            const int rc = after_step(due);
            if (rc != MAVI_OK) return rc;
        }
    } else {
        // only Verlet steps carry forces from one step to the next; the
        // two-step graph replays from state copy 0 and holds no invasion check
        const bool rings = s->P.potential == MAVI_POT_RINGS || is_active(s);
        const bool graphs = s->use_graphs && !s->prof && nsteps >= 4 && !s->var;
        int64_t k = 0;
        while (k < nsteps) {
            // (a rings or self-propelled graph starts from the keys its
            // update passes leave)
            const bool keys_ok = !mavi::update_fuses_bin(s) || s->keys_ready;
            if (graphs && k + 2 <= nsteps && s->cur == 0 && (rings || s->f_valid) && keys_ok &&
                !invasions_due(s, 1)) {
                HIPCHK(ctx, ensure_graph(s, st));
                HIPCHK(ctx, hipGraphLaunch(s->graph2, st));
                if (s->inv_sched) s->inv_since += 2;
                k += 2;
                continue;
            }
            const bool due = invasions_due(s, 0);
            const int rc0 = before_step();
            if (rc0 != MAVI_OK) return rc0;
            HIPCHK(ctx, launch_step(s, st, s->f_valid));
            const int rc = after_step(due);
            if (rc != MAVI_OK) return rc;
            k++;
        }
    }
    int rc = sync_staging(s, true);
    if (rc != MAVI_OK) return rc;
    return check_flags(s);
}

int mavi_step(mavi_sys *s, int64_t nsteps) {
    if (!s) return MAVI_E_BADARG;
    return steps_impl(s, nsteps, nullptr);
}

int mavi_rings_step(mavi_sys *s, int64_t nsteps, const double *noise) {
    if (!s) return MAVI_E_BADARG;
    if (s->P.potential != MAVI_POT_RINGS) return set_err(s->ctx, MAVI_E_BADARG, "not a rings system");
    if (!noise) {
        // zero noise: stage zeros
        mavi_ctx *ctx = s->ctx;
        std::vector<double> z((size_t)(nsteps * s->P.nrings), 0.0);
        (void)ctx;
        return steps_impl(s, nsteps, z.data());
    }
    return steps_impl(s, nsteps, noise);
}

int mavi_sp_step(mavi_sys *s, int64_t nsteps, const double *noise) {
    if (!s) return MAVI_E_BADARG;
    if (!is_active(s)) return set_err(s->ctx, MAVI_E_BADARG, "not a self-propelled (Szabo / RTP) system");
    return steps_impl(s, nsteps, noise);
}

int mavi_rings_neighbors_cfg(mavi_sys *s, int enable, double tol, int only_rings) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (s->P.potential != MAVI_POT_RINGS) return set_err(ctx, MAVI_E_BADARG, "not a rings system");
    if (enable && !(tol > 0.0)) return set_err(ctx, MAVI_E_BADARG, "tol must be > 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (enable && !s->neigh_cnt) {
        HIPCHK(ctx, dalloc(&s->neigh_cnt, s->P.n, st));
        HIPCHK(ctx, dalloc(&s->neigh_list, (int64_t)s->P.n * 15, st));
    }
    if (enable) {
        // max_dist * tol with max_dist = 2 particle_radius(interaction) = dist_eq
        for (auto &I : s->host_inter) I.nb2 = exact_lt_threshold(2.0 * (I.dist_eq / 2.0) * tol);
        HIPCHK(ctx, hipMemcpyAsync(s->inter, s->host_inter.data(), sizeof(RingInter) * s->host_inter.size(),
                                   hipMemcpyHostToDevice, st));
        HIPCHK(ctx, mavi::launch_fill_i32(s->neigh_cnt, 0, s->P.n, st));
        HIPCHK(ctx, mavi::launch_fill_i32(s->neigh_list, -1, (int64_t)s->P.n * 15, st));
    }
    HIPCHK(ctx, hipStreamSynchronize(st));
    const bool changed = s->neigh_on != (enable != 0) || s->neigh_only_rings != (only_rings != 0);
    s->neigh_on = enable != 0;
    s->neigh_only_rings = only_rings != 0;
    s->neigh_tol = tol;
    if (changed && s->graph2) {  // the captured graph holds the old launch sequence
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    return MAVI_OK;
}

int mavi_rings_neighbors_get(mavi_sys *s, int64_t *counts, int64_t *lists) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (!s->neigh_on) return set_err(ctx, MAVI_E_BADARG, "contact neighbours are not enabled");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t n = s->P.n;
    std::vector<int32_t> c(n), l(counts && lists ? n * 15 : 0);
    HIPCHK(ctx, hipMemcpy(c.data(), s->neigh_cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (lists) HIPCHK(ctx, hipMemcpy(l.data(), s->neigh_list, sizeof(int32_t) * n * 15, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
        // neigh_clean! empties every list; a variable population fills the active ones
        const bool act = !s->var || s->vmask[i / s->P.nmax];
        if (counts) counts[i] = act ? c[i] : 0;
        if (lists)
            for (int k = 0; k < 15; k++) lists[i * 15 + k] = act ? l[i * 15 + k] : -1;
    }
    return MAVI_OK;
}

int mavi_rings_invasions_cfg(mavi_sys *s, int scheduled, int64_t steps_to_update, int32_t r_cols, int32_t r_rows,
                             int32_t r_nc) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (s->P.potential != MAVI_POT_RINGS) return set_err(ctx, MAVI_E_BADARG, "not a rings system");
    if (r_cols < 0 || r_rows < 0 || (r_cols == 0) != (r_rows == 0))
        return set_err(ctx, MAVI_E_BADARG, "ring chunk grid: give both r_cols and r_rows (0, 0: no grid)");
    if (r_cols > 0 && r_nc < 1) return set_err(ctx, MAVI_E_BADARG, "ring chunk capacity must be >= 1");
    if (r_cols > 0 && s->P.periodic && (r_cols < 3 || r_rows < 3))
        return set_err(ctx, MAVI_E_UNSUPPORTED,
                       "periodic ring chunk grids need >= 3 chunks per side (the reference stencil repeats "
                       "neighbours below that)");
    if (r_cols > 0 && (int64_t)r_cols * r_rows * r_nc > INT32_MAX)
        return set_err(ctx, MAVI_E_BADARG, "ring chunk grid too large");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int32_t nr = s->P.nrings;
    if (!s->inv_box) {
        HIPCHK(ctx, dalloc(&s->inv_box, nr, st));
        HIPCHK(ctx, dalloc(&s->inv_chunk, nr, st));
        HIPCHK(ctx, dalloc(&s->inv_n, 1, st));
        const int64_t cap = s->P.n > 1024 ? s->P.n : 1024;
        HIPCHK(ctx, dalloc(&s->inv_rec, 3 * cap, st));
        s->inv_cap = cap;
    }
    const int64_t ncell = (int64_t)r_cols * r_rows;
    if (r_cols > 0 && ncell > (int64_t)s->inv_cnt_cap) {
        dfree(s->inv_cnt);
        s->inv_cnt_cap = 0;
        HIPCHK(ctx, dalloc(&s->inv_cnt, ncell, st));
        s->inv_cnt_cap = (int32_t)ncell;
    }
    if (r_cols > 0 && ncell * r_nc > (int64_t)s->inv_cell_cap) {
        dfree(s->inv_cell);
        s->inv_cell_cap = 0;
        HIPCHK(ctx, dalloc(&s->inv_cell, ncell * r_nc, st));
        s->inv_cell_cap = (int32_t)(ncell * r_nc);
    }
    InvGrid G{};
    G.rows = r_rows;
    G.cols = r_cols;
    G.nc = r_nc;
    if (r_cols > 0) {  // chunk_length / chunk_height of the ring Chunks, src/chunks.jl:29-30
        G.cl = s->P.L / (double)r_cols;
        G.ch = s->P.H / (double)r_rows;
    }
    s->inv_grid = G;
    s->inv_sched = scheduled != 0;
    s->inv_every = steps_to_update;
    s->inv_since = 0;  // last_check = num_steps (src/rings/rings.jl:276-278)
    s->inv_count = 0;  // InvasionsInfo(): an empty list
    // the constructor's update_invasions! checks at once only when
    // steps_to_update <= 0
    if (s->inv_sched && s->has_state && steps_to_update <= 0) return run_invasions(s);
    return MAVI_OK;
}

int mavi_rings_invasions_check(mavi_sys *s) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (!s->inv_box) return set_err(ctx, MAVI_E_BADARG, "invasions are not configured (mavi_rings_invasions_cfg)");
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "no state uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return run_invasions(s);
}

int mavi_rings_invasions_get(mavi_sys *s, int64_t *count, int64_t *records, int64_t cap) {
    if (!s || !count) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (!s->inv_box) return set_err(ctx, MAVI_E_BADARG, "invasions are not configured (mavi_rings_invasions_cfg)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t n = s->inv_count;
    *count = n;
    if (!records || n == 0) return MAVI_OK;
    std::vector<int32_t> r((size_t)(3 * n));
    HIPCHK(ctx, hipMemcpy(r.data(), s->inv_rec, sizeof(int32_t) * 3 * n, hipMemcpyDeviceToHost));
    std::vector<int64_t> order((size_t)n);
    for (int64_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        for (int k = 0; k < 3; k++)
            if (r[3 * a + k] != r[3 * b + k]) return r[3 * a + k] < r[3 * b + k];
        return false;
    });
    const int64_t m = n < cap ? n : cap;
    for (int64_t i = 0; i < m; i++)
        for (int k = 0; k < 3; k++) records[3 * i + k] = r[3 * order[i] + k];
    return MAVI_OK;
}

int mavi_rings_var_cfg(mavi_sys *s, const int32_t *mask, int32_t nsrc, const mavi_ring_source *src) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (s->P.potential != MAVI_POT_RINGS) return set_err(ctx, MAVI_E_BADARG, "not a rings system");
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "upload the state before configuring its population");
    if (s->var) return set_err(ctx, MAVI_E_BADARG, "the population is already variable");
    if (nsrc < 0 || (nsrc > 0 && !src)) return set_err(ctx, MAVI_E_BADARG, "bad source list");
    std::vector<HostSource> vs;
    int32_t nsinks = 0, nboxes = 0;
    bool chunked = false;
    const int64_t nm = s->P.nmax, ncell = (int64_t)s->P.ncell;
    for (int32_t q = 0; q < nsrc; q++) {
        const mavi_ring_source &d = src[q];
        HostSource h;
        h.kind = d.kind;
        if (d.kind == MAVI_SRC_SINK) {
            if (d.geom != MAVI_GEOM_RECT && d.geom != MAVI_GEOM_CIRCLE)
                return set_err(ctx, MAVI_E_BADARG, "sink geometry must be a rectangle or a circle");
            if (nsinks >= kMaxSinks) return set_err(ctx, MAVI_E_UNSUPPORTED, "more than 32 sinks");
            h.geom = d.geom;
            for (int k = 0; k < 4; k++) h.g[k] = d.g[k];
            h.sink_bit = nsinks++;
        } else if (d.kind == MAVI_SRC_SOURCE) {
            if (d.nbox < 1 || !d.bbox || !d.spawn) return set_err(ctx, MAVI_E_BADARG, "source without boxes");
            if (nboxes + d.nbox > kMaxBoxes) return set_err(ctx, MAVI_E_UNSUPPORTED, "more than 64 spawn boxes");
            if (d.chunk_off) {
                if (s->P.cols == 0)
                    return set_err(ctx, MAVI_E_BADARG, "ChunksChecker lists given for a system without chunks");
                for (int32_t e = d.chunk_off[0]; e < d.chunk_off[d.nbox]; e++)
                    if (d.chunk_ids[e] < 0 || d.chunk_ids[e] >= ncell)
                        return set_err(ctx, MAVI_E_BADARG, "ChunksChecker chunk index out of range");
                chunked = true;
            }
            h.nbox = d.nbox;
            h.random_pol = d.random_pol;
            h.pad = d.pad;
            h.pol = d.pol;
            h.box0 = nboxes;
            h.spawn.assign(d.spawn, d.spawn + (size_t)d.nbox * 2 * nm);
            nboxes += d.nbox;
        } else {
            return set_err(ctx, MAVI_E_BADARG, "unknown source kind");
        }
        vs.push_back(std::move(h));
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int32_t nr = s->P.nrings;
    HIPCHK(ctx, dalloc(&s->rmask, nr, st));
    HIPCHK(ctx, dalloc(&s->dsinks, nsinks > 0 ? nsinks : 1, st));
    HIPCHK(ctx, dalloc(&s->dboxes, nboxes > 0 ? nboxes : 1, st));
    {
        char *vrb = nullptr;
        HIPCHK(ctx, dalloc(&vrb, (int64_t)kVarHeader + sizeof(int32_t) * 2 * (int64_t)nr, st));
        s->vrb = vrb;
        s->sink_n = (uint32_t *)vrb;
        s->nonempty = (unsigned long long *)(vrb + 8);
        s->sink_list = (int32_t *)(vrb + kVarHeader);
    }
    s->vbits.assign(nr, 0u);
    s->vadded.assign(nr, 0);
    s->vacm.assign(2 * (size_t)nr, 0.0);
    if (chunked) HIPCHK(ctx, dalloc(&s->box_mask, ncell, st));
    s->vstage_bytes = var_stage_bytes(s);
    HIPCHK(ctx, hipHostMalloc(&s->vstage, s->vstage_bytes, hipHostMallocDefault));
    // device tables (synchronous copies: the host vectors are temporaries)
    std::vector<DevSink> ds;
    std::vector<DevBox> db;
    std::vector<unsigned long long> bm(chunked ? (size_t)ncell : 0, 0ull);
    for (int32_t q = 0; q < nsrc; q++) {
        const mavi_ring_source &d = src[q];
        if (d.kind == MAVI_SRC_SINK) {
            DevSink k{};
            k.geom = d.geom;
            for (int j = 0; j < 4; j++) k.g[j] = d.g[j];
            ds.push_back(k);
            continue;
        }
        for (int32_t b = 0; b < d.nbox; b++) {
            DevBox k{};
            k.x = d.bbox[4 * b];
            k.y = d.bbox[4 * b + 1];
            k.len = d.bbox[4 * b + 2];
            k.h = d.bbox[4 * b + 3];
            k.pad = d.pad;
            k.chunked = d.chunk_off ? 1 : 0;
            const int32_t gb = (int32_t)db.size();
            db.push_back(k);
            if (d.chunk_off)
                for (int32_t e = d.chunk_off[b]; e < d.chunk_off[b + 1]; e++) bm[d.chunk_ids[e]] |= 1ull << gb;
        }
    }
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (!ds.empty()) HIPCHK(ctx, hipMemcpy(s->dsinks, ds.data(), sizeof(DevSink) * ds.size(), hipMemcpyHostToDevice));
    if (!db.empty()) HIPCHK(ctx, hipMemcpy(s->dboxes, db.data(), sizeof(DevBox) * db.size(), hipMemcpyHostToDevice));
    if (chunked)
        HIPCHK(ctx, hipMemcpy(s->box_mask, bm.data(), sizeof(unsigned long long) * bm.size(), hipMemcpyHostToDevice));
    // VarRingsIds(mask, ids = 1:n, uids = 1:n), then update_ids!
    s->vmask.assign(nr, 1);
    if (mask)
        for (int32_t r = 0; r < nr; r++) s->vmask[r] = mask[r] ? 1 : 0;
    s->vids.resize(nr);
    s->vuids.resize(nr);
    for (int32_t r = 0; r < nr; r++) s->vids[r] = s->vuids[r] = r + 1;
    s->vsrc = std::move(vs);
    s->nsinks = nsinks;
    s->nboxes = nboxes;
    s->pol_draws.clear();
    s->pol_used = 0;
    s->pol_stream = 0;
    s->var = true;
    s->P.rmask = s->rmask;
    if (s->graph2) {
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    std::vector<int32_t> active;
    var_update_ids(s, active);
    // rebuild the state from the id-order staging with the active set, zero
    // the inactive rings' centres of mass, then the constructor's first
    // update_chunks!
    int rc = var_apply(s, {}, active);
    if (rc != MAVI_OK) return rc;
    HIPCHK(ctx, mavi::launch_var_clear(s, st));
    HIPCHK(ctx, mavi::launch_bin(s, st, false, false, false));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return check_flags(s);
}

int mavi_rings_pol_draws(mavi_sys *s, const double *u, int64_t n) {
    if (!s || n < 0 || (n > 0 && !u)) return MAVI_E_BADARG;
    if (!s->var) return set_err(s->ctx, MAVI_E_BADARG, "not a variable ring population (mavi_rings_var_cfg)");
    s->pol_draws.insert(s->pol_draws.end(), u, u + n);
    return MAVI_OK;
}

int mavi_rings_pol_stream_set(mavi_sys *s, int64_t position) {
    if (!s || position < 0) return MAVI_E_BADARG;
    if (!s->var) return set_err(s->ctx, MAVI_E_BADARG, "not a variable ring population (mavi_rings_var_cfg)");
    s->pol_stream = position;
    return MAVI_OK;
}

int mavi_rings_var_get(mavi_sys *s, int32_t *mask, int64_t *ids, int64_t *uids, int64_t *num_active,
                       int64_t *pol_used, int64_t *pol_stream) {
    if (!s) return MAVI_E_BADARG;
    if (!s->var) return set_err(s->ctx, MAVI_E_BADARG, "not a variable ring population (mavi_rings_var_cfg)");
    const int32_t nr = s->P.nrings;
    for (int32_t r = 0; r < nr; r++) {
        if (mask) mask[r] = s->vmask[r];
        if (ids) ids[r] = s->vids[r];
        if (uids) uids[r] = s->vuids[r];
    }
    if (num_active) *num_active = s->vnum_active;
    if (pol_used) *pol_used = s->pol_used;
    if (pol_stream) *pol_stream = s->pol_stream;
    return MAVI_OK;
}

int mavi_rings_info(mavi_sys *s, double *areas, double *cms, double *continuous_pos) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (s->P.potential != MAVI_POT_RINGS) return set_err(ctx, MAVI_E_BADARG, "not a rings system");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (areas) HIPCHK(ctx, hipMemcpy(areas, s->areas, sizeof(double) * s->P.nrings, hipMemcpyDeviceToHost));
    if (cms) HIPCHK(ctx, hipMemcpy(cms, s->cms, sizeof(double2) * s->P.nrings, hipMemcpyDeviceToHost));
    if (continuous_pos)
        HIPCHK(ctx, hipMemcpy(continuous_pos, s->cpos, sizeof(double2) * s->P.n, hipMemcpyDeviceToHost));
    return MAVI_OK;
}

int mavi_observables(mavi_sys *s, double *ke, double *pe) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (is_active(s)) return set_err(ctx, MAVI_E_BADARG, "a self-propelled state has no velocities");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nb = (s->P.n + 255) / 256;
    if (s->red_cap < 2 * nb) {
        dfree(s->red);
        s->red_cap = 0;
        HIPCHK(ctx, dalloc(&s->red, 2 * nb, ctx->stream));
        s->red_cap = 2 * nb;
    }
    HIPCHK(ctx, mavi::launch_observables(s, s->red, ctx->stream));
    std::vector<double> h(2 * nb);
    HIPCHK(ctx, hipMemcpyAsync(h.data(), s->red, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double k = 0.0, p = 0.0;
    for (int b = 0; b < nb; b++) {
        k += h[2 * b];
        p += h[2 * b + 1];
    }
    if (ke) *ke = k / 2.0;
    if (pe) *pe = (s->P.potential == MAVI_POT_LJ) ? p * (4.0 * s->P.p1) : NAN;
    return MAVI_OK;
}

int mavi_run_record(mavi_sys *s, int64_t nsteps, double *ke, double *pe) {
    if (!s || !ke || s->P.potential == MAVI_POT_RINGS || is_active(s)) return MAVI_E_BADARG;
    for (int64_t k = 0; k < nsteps; k++) {
        int st = mavi_step(s, 1);
        if (st != MAVI_OK) return st;
        double p = 0.0;
        st = mavi_observables(s, &ke[k], &p);
        if (st != MAVI_OK) return st;
        if (pe) pe[k] = p;
    }
    return MAVI_OK;
}

int mavi_prof_enable(mavi_sys *s, int on) {
    if (!s) return MAVI_E_BADARG;
    s->prof = on != 0;
    for (auto &p : s->profs) p = ProfSlot{};
    return MAVI_OK;
}

int mavi_prof_get(mavi_sys *s, int which, double *ms_total, int64_t *launches) {
    if (!s || which < 0 || which >= MAVI_PROF_NCLASS) return MAVI_E_BADARG;
    if (ms_total) *ms_total = s->profs[which].ms;
    if (launches) *launches = s->profs[which].launches;
    return MAVI_OK;
}

int mavi_nlist_info(mavi_sys *s, int64_t *info) {
    if (!s || !info) return MAVI_E_BADARG;
    const NList *N = s->nl;
    info[0] = N ? 1 : 0;
    info[1] = (N && N->on && !N->disabled) ? 1 : 0;
    info[2] = N ? N->rebuilds : 0;
    info[3] = N ? N->rollbacks : 0;
    return MAVI_OK;
}

int mavi_sys_set_tile(mavi_sys *s, int32_t workgroup) {
    if (s) s->f_valid = false;
    if (!s) return MAVI_E_BADARG;
    if (workgroup == 0) workgroup = 64;
    if (!is_pow2_wg(workgroup)) return set_err(s->ctx, MAVI_E_BADARG, "workgroup must be 64..1024, power of two");
    if (workgroup != s->wg && s->graph2) {
        (void)hipStreamSynchronize(s->ctx->stream);
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    s->wg = workgroup;
    return MAVI_OK;
}

int mavi_sys_noise_get(mavi_sys *s, uint64_t *seed, uint64_t *counter) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    unsigned long long c = 0;
    HIPCHK(ctx, hipMemcpyAsync(&c, s->counter, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (seed) *seed = s->seed;
    if (counter) *counter = c;
    return MAVI_OK;
}

int mavi_sys_noise_set(mavi_sys *s, uint64_t seed, uint64_t counter) {
    if (!s) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned long long c = counter;
    HIPCHK(ctx, hipMemcpy(s->counter, &c, sizeof(c), hipMemcpyHostToDevice));
    if (seed != s->seed && s->graph2) {  // the captured graph holds the seed
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    s->seed = seed;
    return MAVI_OK;
}

int mavi_sys_set_option(mavi_sys *s, int32_t option, int32_t value) {
    if (!s) return MAVI_E_BADARG;
    bool changed = false;
    switch (option) {
        case MAVI_OPT_TWO_PHASE:
            changed = s->two_phase != (value != 0);
            s->two_phase = value != 0;
            break;
        case MAVI_OPT_GRAPHS:
            s->use_graphs = value != 0;
            break;
        case MAVI_OPT_LDS_TILES:
            changed = s->lds_tiles != (value != 0);
            s->lds_tiles = value != 0;
            break;
        case MAVI_OPT_REUSE_FORCES:
            changed = s->reuse != (value != 0);
            s->reuse = value != 0;
            break;
        case MAVI_OPT_NLIST:
            if (s->nl) {
                s->nl->on = value != 0;
                s->nl->valid = false;
            }
            changed = true;
            break;
        case MAVI_OPT_NLIST_SKIN:
            if (value < 0 || value > 2000) return set_err(s->ctx, MAVI_E_BADARG, "skin: 0 (default) .. 2000 thousandths");
            if (s->nl) {
                nl_drop_graph(s);
                mavi::nl_set_skin(s, value == 0 ? 0.15 : value / 1000.0);
            }
            break;
        case MAVI_OPT_NLIST_BLOCK:
            if (value < 0 || value > 4096) return set_err(s->ctx, MAVI_E_BADARG, "block: 0 (default) .. 4096 steps");
            if (s->nl) {
                s->nl->block = (int32_t)value;
                s->nl->block_now = value == 0 ? kNlBlockCap : (int32_t)value;
            }
            break;
        default:
            return set_err(s->ctx, MAVI_E_BADARG, "unknown option");
    }
    if (changed) s->f_valid = false;
    if (changed) nl_drop_graph(s);
    if (changed && s->graph2) {  // the captured graph holds the old kernels
        (void)hipStreamSynchronize(s->ctx->stream);
        (void)hipGraphExecDestroy(s->graph2);
        s->graph2 = nullptr;
    }
    return MAVI_OK;
}

int mavi_autotune(mavi_sys *s, int32_t trial_steps, double *best_us, int32_t *shape) {
    if (!s || trial_steps <= 0) return MAVI_E_BADARG;
    mavi_ctx *ctx = s->ctx;
    if (!s->has_state) return set_err(ctx, MAVI_E_BADARG, "no state uploaded");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // snapshot of the id-order state; every trial starts from it
    int rc = sync_staging(s, true);
    if (rc != MAVI_OK) return rc;
    const size_t bytes = sizeof(double2) * (size_t)s->P.n;
    double2 *pos0 = nullptr, *vel0 = nullptr;
    double *pol0 = nullptr;
    HIPCHK(ctx, dalloc(&pos0, s->P.n, st));
    hipError_t e = hipMemcpyAsync(pos0, s->pos_id, bytes, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && s->vel_id) {
        e = dalloc(&vel0, s->P.n, st);
        if (e == hipSuccess) e = hipMemcpyAsync(vel0, s->vel_id, bytes, hipMemcpyDeviceToDevice, st);
    }
    if (e == hipSuccess && s->pol) {
        e = dalloc(&pol0, s->P.nrings, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(pol0, s->pol, sizeof(double) * s->P.nrings, hipMemcpyDeviceToDevice, st);
    }
    const unsigned long long *counter0 = nullptr;
    (void)counter0;
    auto restore = [&]() -> hipError_t {
        hipError_t r = hipMemcpyAsync(s->pos_id, pos0, bytes, hipMemcpyDeviceToDevice, st);
        if (r == hipSuccess && vel0) r = hipMemcpyAsync(s->vel_id, vel0, bytes, hipMemcpyDeviceToDevice, st);
        if (r == hipSuccess && pol0)
            r = hipMemcpyAsync(s->pol, pol0, sizeof(double) * s->P.nrings, hipMemcpyDeviceToDevice, st);
        s->cur = 0;
        s->f_valid = false;
        if (r == hipSuccess) r = mavi::drop_keys(s, st);
        if (r == hipSuccess) r = mavi::launch_iota_state(s, st);
        if (r == hipSuccess && s->P.potential == MAVI_POT_RINGS) r = mavi::launch_rings_init(s, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        return r;
    };
    const int wgs[4] = {64, 128, 256, 512};
    double best = 1e300;
    int bw = s->wg, bl = s->lds_tiles, bt = s->two_phase;
    const bool saved_prof = s->prof;
    s->prof = false;
    for (int lds = 0; lds < 2 && e == hipSuccess; lds++) {
        if (lds && s->P.cols == 0) continue;
        for (int tp = 0; tp < 2 && e == hipSuccess; tp++) {
            for (int wi = 0; wi < 4 && e == hipSuccess; wi++) {
                if (lds && wi > 0) continue;  // the LDS form has a fixed workgroup
                if (mavi_sys_set_tile(s, wgs[wi]) != MAVI_OK) continue;
                mavi_sys_set_option(s, MAVI_OPT_LDS_TILES, lds);
                mavi_sys_set_option(s, MAVI_OPT_TWO_PHASE, tp);
                if ((e = restore()) != hipSuccess) break;
                if (steps_impl(s, 2, nullptr) != MAVI_OK) continue;  // warm-up (and graph capture)
                hipEvent_t a = s->ev0, b = s->ev1;
                e = hipEventRecord(a, st);
                if (e == hipSuccess && steps_impl(s, trial_steps, nullptr) != MAVI_OK) continue;
                if (e == hipSuccess) e = hipEventRecord(b, st);
                if (e == hipSuccess) e = hipEventSynchronize(b);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
                const double us = 1e3 * ms / trial_steps;
                if (e == hipSuccess && us < best) {
                    best = us;
                    bw = wgs[wi];
                    bl = lds;
                    bt = tp;
                }
            }
        }
    }
    s->prof = saved_prof;
    mavi_sys_set_tile(s, bw);
    mavi_sys_set_option(s, MAVI_OPT_LDS_TILES, bl);
    mavi_sys_set_option(s, MAVI_OPT_TWO_PHASE, bt);
    if (e == hipSuccess) e = restore();
    dfree(pos0);
    dfree(vel0);
    dfree(pol0);
    HIPCHK(ctx, e);
    if (best_us) *best_us = best;
    if (shape) {
        shape[0] = bw;
        shape[1] = bl;
        shape[2] = bt;
    }
    return MAVI_OK;
}

}  // extern "C"
