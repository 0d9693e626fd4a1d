/*
 * mavi_oracle.h -- CPU restatement of Mavi.jl's step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library is compared against; it is never linked into, loaded by or called
 * from the product path (libmavi_hip.so / the mavi_hip Python package).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function restates a Julia function of the reference (Mavi.jl v1.0.1);
 * the citation is given next to each declaration as path:line relative to the
 * reference repository root.  Particle ids are 1-based, as in Julia.  Host
 * positions/velocities/forces are interleaved x0,y0,x1,y1,... exactly like
 * Julia's Vector{SVector{2,Float64}}.
 *
 * Pinning (what checks each function independently of its own formula) is
 * listed in the header comment of mavi_oracle.c.
 */
#ifndef MAVI_ORACLE_H
#define MAVI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_KIND_LJ = 0, OR_KIND_HARMTRUNC = 1, OR_KIND_RINGS = 2, OR_KIND_SZABO = 3, OR_KIND_RTP = 4 };
enum { OR_OK = 0, OR_E_CELL_OVERFLOW = 1, OR_E_OUT_OF_BOX = 2, OR_E_BADARG = 5 };

/* One (wall, geometry) pair of a SpaceCfg, applied in list order (ManyWalls,
 * src/configs.jl:242-244, 289-301; src/integration.jl:257-263, 398-403). */
enum {
    OR_WK_RIGID_RECT = 0,        /* walls! src/integration.jl:271-285 */
    OR_WK_PERIODIC_RECT = 1,     /* walls! :309-324 */
    OR_WK_RIGID_CIRCLE = 2,      /* walls! :287-307 */
    OR_WK_SLIPPERY_LINES = 3,    /* walls! :327-375 */
    OR_WK_SLIPPERY_CIRCLE = 4,   /* walls! :377-396 */
    OR_WK_POTENTIAL_CIRCLE = 5,  /* calc_walls_forces! :230-240 */
    OR_WK_POTENTIAL_LINES = 6    /* calc_walls_forces! :242-255 */
};
enum { OR_WMODE_OUTSIDE = 0, OR_WMODE_INSIDE = 1, OR_WMODE_REPULSION = 2 }; /* src/configs.jl:255-261 */
typedef struct {
    int32_t kind, mode;   /* OR_WK_*, OR_WMODE_* (potential walls) */
    int32_t pot_kind;     /* potential walls: OR_KIND_LJ or OR_KIND_HARMTRUNC */
    int32_t nlines;       /* LinesCfg: number of segments */
    double geom[4];       /* rectangle: bl_x, bl_y, length, height; circle: center_x, center_y, radius */
    double pot[4];        /* potential walls: LJ sigma, epsilon; HT k_rep, k_atr, dist_eq, dist_max */
    const double *lines;  /* LinesCfg: nlines x (p1x, p1y, p2x, p2y) */
} or_wall;

/* One system.  Mirrors what Mavi's System holds for the hot path:
 *   space_cfg  (src/configs.jl:285-288) -> bl, size, periodic
 *   dynamic_cfg (src/configs.jl:341-397, src/rings/configs.jl:95-107) -> kind, p, ring_*
 *   int_cfg    (src/configs.jl:482-487) -> dt, cols/rows (ChunksCfg)
 *   Chunks     (src/chunks.jl:10-40)    -> cols, rows, cl, ch, nc
 */
/* Variable ring population: VarRingsIds (src/rings/states.jl:20-39,
 * 173-223) and the RingsInfo fields that persist between steps.  Owned by
 * the caller, updated in place by or_rings_var_init / or_rings_steps. */
typedef struct {
    int32_t *mask;        /* nrings, 1 = active */
    int64_t *ids;         /* nrings, VarRingsIds.ids, 1-based; entries past
                             num_active keep what an earlier update left */
    int64_t *uids;        /* nrings */
    int64_t num_active;
    double *cms;          /* 2*nrings, RingsInfo.cms (stale for inactive rings) */
    double *cpos;         /* 2*nmax*nrings, RingsInfo.continuos_pos */
    int64_t *cp, *counts; /* the particle chunks of the last update_chunks!
                             (nc*rows*cols, rows*cols), read by ChunksChecker */
    const double *pol_draws;   /* uniform draws consumed by spawn_pol = :random */
    int64_t npol_draws, pol_used;
} or_var_rings;

/* A source or a sink of rings (src/rings/sources.jl), processed in list
 * order by update_sources! (src/rings/integration.jl:353-358). */
enum { OR_SRC_SOURCE = 0, OR_SRC_SINK = 1 };
enum { OR_GEOM_RECT = 0, OR_GEOM_CIRCLE = 1 };
typedef struct {
    int32_t kind;          /* OR_SRC_* */
    int32_t geom;          /* sink geometry: OR_GEOM_RECT (blx, bly, length, height) or OR_GEOM_CIRCLE (cx, cy, r) */
    double g[4];
    int32_t nbox;          /* source: Source.bbox boxes */
    int32_t random_pol;    /* 1: spawn_pol = :random (2 pi u per empty box), 0: the fixed pol below */
    double pad;            /* SourceCfg.pad */
    double pol;
    const double *bbox;    /* nbox x 4: bottom_left x, y, length, height */
    const double *spawn;   /* nbox x nmax x 2: Source.bbox_spawn_pos */
    const int32_t *chunk_off;  /* nbox+1 offsets into chunk_ids; NULL: PosChecker */
    const int32_t *chunk_ids;  /* ChunksChecker.ids as 0-based row0*cols + col0 */
} or_source;

typedef struct {
    int64_t n;          /* particle slots (length(state.pos)) */
    double bl[2];       /* geometry_cfg.bottom_left */
    double size[2];     /* geometry_cfg.length, geometry_cfg.height */
    int32_t periodic;   /* 1: PeriodicWalls, 0: RigidWalls */
    int32_t kind;       /* OR_KIND_* */
    double p[4];        /* LJ: sigma, epsilon, rcut (<=0: no cutoff, reference semantics)
                           HT: k_rep, k_atr, dist_eq, dist_max
                           Szabo: k_rep, k_adh, r_eq, r_max
                           RTP: sigma, epsilon, cutoff 2^(1/6) sigma */
    double dt;
    int32_t cols, rows; /* ChunksCfg; cols == 0 -> chunks = nothing (all pairs) */
    int32_t nc;         /* chunk capacity (src/chunks.jl:32-34) */
    int32_t nthreads;   /* rings steps: threads of calc_forces! (Threaded mode) and the per-ring loops; <= 1: one */
    double cl, ch;      /* chunk_length, chunk_height (src/chunks.jl:29-30) */
    double radius;      /* particle_radius(dynamic_cfg) (src/configs.jl:418-421) */
    /* Rings (kind == OR_KIND_RINGS) */
    int32_t nmax, nrings, ntypes, pad1;
    const int32_t *types;    /* per ring, 1-based type; NULL -> untyped (one type) */
    const int32_t *np_type;  /* particles per ring, per type */
    const double *ring_par;  /* per type, 8 doubles: p0, relax_time, vo, mobility,
                                rot_diff, k_area, k_spring, l_spring */
    const double *inter;     /* ntypes x ntypes HarmTrunc params, 4 doubles each,
                                index (t1-1) + ntypes*(t2-1) (Julia column-major) */
    /* Self-propelled update (kind SZABO: vo, mobility, relax_time, rot_diff;
       kind RTP: vo, tumble_rate, -, -) */
    double ap[4];
    /* Walls list (nwalls > 0): replaces `periodic`-selected rigid/periodic
       rectangle walls; the first entry is the main wall (calc_diff, stencil) */
    int32_t nwalls, pad2;
    const or_wall *walls;
    /* Rings with a variable population (NULL: FixRingsIds) and their
       sources and sinks */
    or_var_rings *var;
    int32_t nsrc, pad3;
    const or_source *src;
} or_sys;

/* Julia Base `div(a::Float64, b::Float64)` (RoundToZero): round((a - rem(a,b))/b). */
double or_jl_div(double a, double b);

/* Chunk capacity nc, src/chunks.jl:32-34. */
int64_t or_chunk_capacity(double chunk_length, double chunk_height, double particle_r);

/* Cell of one position, src/chunks.jl:120-142 (1-based row, col). */
int or_cell(const or_sys *s, double x, double y, int64_t *row, int64_t *col);

/* Active particle ids in ascending order (get_particles_ids):
 * src/states.jl:186-228 (all ids), src/rings/states.jl:46-62 (FixRingsIds).
 * Returns the count; ids (1-based) written when ids != NULL. */
int64_t or_particle_ids(const or_sys *s, int64_t *ids);

/* update_chunks!, src/chunks.jl:150-161 with src/integration.jl:54-59 and
 * src/rings/integration.jl:18-23.  chunk_particles is (nc, rows, cols)
 * column-major, counts is (rows, cols) column-major, ids 1-based. */
int or_bin(const or_sys *s, const double *pos, int64_t *chunk_particles, int64_t *counts);

/* get_neighbors, src/chunks.jl:61-118.  For cell (r,c) (1-based) the list is
 * stored at k = (c-1)*rows + (r-1): nnb[k] entries, nb[(k*4 + e)*2 + {0,1}] = (row, col). */
void or_neighbors(int32_t rows, int32_t cols, int32_t periodic, int64_t *nb, int32_t *nnb);

/* calc_diff, src/integration.jl:37-48. */
void or_calc_diff(const or_sys *s, const double *r1, const double *r2, double *out);

/* calc_interaction(i, j) for the configured potential:
 * src/integration.jl:62-66 + src/configs.jl:321-323, 354-368, 389-397 (LJ/HT),
 * src/rings/integration.jl:32-77 (rings).  Ids 1-based. */
void or_pair_force(const or_sys *s, const double *pos, int64_t i, int64_t j, double *f);

/* calc_forces!: Sequencial chunk loop (src/integration.jl:112-157) when
 * s->cols > 0, all-pairs (src/integration.jl:197-224) otherwise.  Zeroes
 * forces first (clean_forces!, src/systems.jl:119-123).  absum (nullable)
 * receives per particle sum_j |f_ij| (both components' magnitudes), the
 * scale used by the parity metric. */
int or_calc_forces(const or_sys *s, const double *pos, const int64_t *chunk_particles,
                   const int64_t *counts, double *forces, double *absum);

/* Threaded calc_forces!, src/integration.jl:159-194: static column split over
 * nthreads, per-thread force slices, slices summed in order. */
int or_calc_forces_threaded(const or_sys *s, const double *pos, const int64_t *chunk_particles,
                            const int64_t *counts, double *forces, int nthreads);

/* newton_step! x nsteps, src/integration.jl:415-431, 507-515 with walls!
 * :271-285 / :309-324.  pos/vel updated in place; forces receives
 * get_forces(system) after the last step; ke/pe (nullable) receive
 * kinetic_energy / potential_energy after every step (src/quantities.jl).
 * nthreads > 1 uses the Threaded force loop. */
int or_newton_steps(const or_sys *s, double *pos, double *vel, double *forces,
                    int64_t nsteps, double *ke, double *pe, int nthreads);

/* ParticleNeighbors of a rings system (src/rings/neighbors.jl:117-137) as
 * filled by calc_interaction (src/rings/integration.jl:32-45) during one
 * calc_forces! traversal (chunk loop or all pairs) at positions pos: a
 * visited pair closer than tol * 2 * particle_radius(interaction) (only
 * pairs of different rings when only_rings) is appended to both lists.
 * counts (nslots), lists (15 x nslots, column per particle, 1-based ids in
 * visiting order, -1 padded).  Returns OR_E_CELL_OVERFLOW past 15 (the
 * BoundsError of neigh_update_data!). */
int or_rings_neighbors(const or_sys *s, const double *pos, double tol, int only_rings, int64_t *counts,
                       int64_t *lists);

/* point_line_intersect, src/rings/integration.jl:379-419: 1 if the ray from
 * p towards +x crosses the segment (a, b) under the reference's strict tests. */
int or_point_line_intersect(const double *p, const double *a, const double *b);
/* check_invasions!(system), src/rings/integration.jl:379-509, at the ring
 * points rpts (ring_points: info.continuos_pos for a periodic main wall,
 * state.rings_pos otherwise, 2*nmax*nrings) and centres of mass cms
 * (2*nrings).  rcols == 0: all ring pairs (r_chunks = nothing); otherwise
 * r_chunks of rrows x rcols chunks of capacity rnc over the chunk space.
 * rec receives up to cap records (invasor, invaded, p_id), 1-based, in the
 * reference's push order; *count the total.  Returns OR_E_OUT_OF_BOX /
 * OR_E_CELL_OVERFLOW where binning a centre of mass throws BoundsError. */
int or_rings_invasions(const or_sys *s, const double *rpts, const double *cms, int32_t rcols, int32_t rrows,
                       int64_t rnc, int64_t *rec, int64_t cap, int64_t *count);

/* calc_walls_forces! for particle i (0-based): adds the potential walls'
 * forces to f[0..1], src/integration.jl:230-255. */
void or_wall_forces(const or_sys *s, const double *pos, int64_t i, double *f);
/* walls! for particle i (0-based) over the walls list, in list order; vel may
 * be NULL when no rigid wall is listed; radius is get_particle_radius. */
void or_walls_apply(const or_sys *s, double *pos, double *vel, int64_t i, double radius);

/* kinetic_energy, src/quantities.jl:12-18. */
double or_kinetic_energy(int64_t n, const double *vel);
/* potential_energy(::LenJonesCfg), src/quantities.jl:46-66 (all pairs, no cutoff). */
double or_potential_energy_lj(const or_sys *s, const double *pos);

/* Rings.  cpos = info.continuos_pos (2*nmax*nrings), areas (nrings), cms (2*nrings). */
/* update_continuos_pos!, src/rings/integration.jl:118-136 (periodic), copy otherwise. */
void or_rings_continuous_pos(const or_sys *s, const double *pos, double *cpos);
/* calc_area, src/rings/integration.jl:103-116 over ring r (1-based). */
double or_rings_area(const or_sys *s, const double *cpos, int32_t ring);
/* forces! (pair + springs + area), src/rings/integration.jl:140-226. */
int or_rings_forces(const or_sys *s, const double *pos, const double *cpos,
                    double *forces, double *areas);
/* The RingsSystem constructor's sequence for a variable population
 * (src/rings/rings.jl:280-284, src/rings/states.jl:106-123): update_ids!
 * from var->mask (ids start as 1:nrings), continuos_pos and cms of the
 * active rings (cms of the others 0), the first particle binning. */
int or_rings_var_init(const or_sys *s, const double *pos);
/* Rings.Integration.step! x nsteps, src/rings/integration.jl:522-543.
 * noise (nullable) holds nsteps*nrings standard normal draws consumed in
 * ring order per step (the randn(system.rng) of :348), indexed by ring slot;
 * NULL means zero noise.  forces/areas/cms (nullable) receive the last
 * step's values.  With s->var, a step also runs update_cms! over the active
 * rings, the sources and sinks in list order, update_ids! and keeps the
 * persistent fields of s->var (continue from or_rings_var_init or from a
 * previous call). */
int or_rings_steps(const or_sys *s, double *pos, double *pol, int64_t nsteps,
                   const double *noise, double *forces, double *areas, double *cms);

/* szabo_step! / rtp_step! x nsteps, src/integration.jl:433-489, 517-535:
 * update_chunks!, calc_forces! (Szabo calc_interaction :68-86, run-and-tumble
 * :88-109), update_szabo! / update_rtp!, periodic walls!.  pos and the
 * polarisation angles ang (n) are updated in place; forces (nullable)
 * receives the last step's forces.  noise (nullable = zero noise / never
 * tumble) replaces the global RNG draws, per step and particle id:
 *   Szabo: noise[step*n + i]           the randn() of :459
 *   RTP:   noise[(step*n + i)*2 + 0]   the rand() of :484 (tumble test)
 *          noise[(step*n + i)*2 + 1]   the rand() of :486 (new angle)
 * (the reference draws the second uniform only when tumbling; here both are
 * given, so a stream of draws cannot be reproduced, only the values). */
int or_sp_steps(const or_sys *s, double *pos, double *ang, int64_t nsteps, const double *noise, double *forces,
                int nthreads);

#ifdef __cplusplus
}
#endif
#endif
