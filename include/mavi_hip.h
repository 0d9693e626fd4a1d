/*
 * mavi_hip.h -- C ABI of libmavi_hip.so, the MI355X (gfx950) step hot path of
 * Mavi.jl.  Plain C types only: host arrays are caller-owned and never kept
 * after a call returns; device buffers, streams and graphs are owned by the
 * library.  Every function returns MAVI_OK (0) or an MAVI_E_* status; the
 * message of the last failure on a context is mavi_last_error(ctx).
 *
 * Host array layout is Julia's Vector{SVector{2,Float64}}: x0,y0,x1,y1,...
 * Particle ids are the slot indices of that vector (0-based here, 1-based in
 * Julia); for rings, slot = (ring-1)*Nmax + p as in to_scalar_idx
 * (src/rings/states.jl:137-140).
 *
 * Each entry point names the reference function whose semantics it
 * reproduces (paths relative to the Mavi.jl repository root).
 */
#ifndef MAVI_HIP_H
#define MAVI_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MAVI_ABI_VERSION 2

/* Status codes. */
#define MAVI_OK 0
#define MAVI_E_CELL_OVERFLOW 1 /* a chunk holds more than nc particles: the
                                  BoundsError of src/chunks.jl:144-146 */
#define MAVI_E_OUT_OF_BOX 2    /* a particle maps outside the chunk grid: the
                                  BoundsError of src/chunks.jl:144 / the throw
                                  of src/systems.jl:76-79 */
#define MAVI_E_NONFINITE 3     /* NaN/Inf position met while binning */
#define MAVI_E_HIP 4           /* HIP runtime failure */
#define MAVI_E_BADARG 5        /* invalid descriptor or argument */
#define MAVI_E_UNSUPPORTED 6   /* configuration outside the GPU path (the
                                  Julia glue keeps the CPU method for it) */

/* SpaceCfg wall types (src/configs.jl:238-239). */
#define MAVI_WALL_RIGID 0
#define MAVI_WALL_PERIODIC 1

/* Built-in DynamicCfg kinds. */
#define MAVI_POT_LJ 0        /* LenJonesCfg, src/configs.jl:380-397 */
#define MAVI_POT_HARMTRUNC 1 /* HarmTruncCfg, src/configs.jl:341-368 */
#define MAVI_POT_RINGS 2     /* RingsCfg + HarmTruncCfg/InteractionMatrix,
                                src/rings/configs.jl:26-164 */
#define MAVI_POT_SZABO 3     /* SzaboCfg, src/configs.jl:399-408, with a
                                SelfPropelledState (src/states.jl:104-125) */
#define MAVI_POT_RTP 4       /* RunTumbleCfg, src/configs.jl:410-415, with a
                                SelfPropelledState */

typedef struct mavi_ctx mavi_ctx;
typedef struct mavi_sys mavi_sys;

/* One (wall, geometry) pair of a SpaceCfg.  A list of them is ManyWalls /
 * ManyGeometries (src/configs.jl:30-32, 242-244, 289-301), applied in list
 * order; its first entry is the main wall (calc_diff and the chunk stencil,
 * src/integration.jl:50-52, src/chunks.jl:27). */
#define MAVI_WK_RIGID_RECT 0        /* RigidWalls + RectangleCfg: walls! src/integration.jl:271-285 */
#define MAVI_WK_PERIODIC_RECT 1     /* PeriodicWalls + RectangleCfg: walls! :309-324 */
#define MAVI_WK_RIGID_CIRCLE 2      /* RigidWalls + CircleCfg: walls! :287-307 */
#define MAVI_WK_SLIPPERY_LINES 3    /* SlipperyWalls + LinesCfg: walls! :327-375 */
#define MAVI_WK_SLIPPERY_CIRCLE 4   /* SlipperyWalls + CircleCfg: walls! :377-396 */
#define MAVI_WK_POTENTIAL_CIRCLE 5  /* PotentialWalls + CircleCfg: calc_walls_forces! :230-240 */
#define MAVI_WK_POTENTIAL_LINES 6   /* PotentialWalls + LinesCfg: calc_walls_forces! :242-255 */
#define MAVI_WMODE_OUTSIDE 0        /* PotentialWallMode, src/configs.jl:255-261 */
#define MAVI_WMODE_INSIDE 1
#define MAVI_WMODE_REPULSION 2
typedef struct {
    int32_t kind;           /* MAVI_WK_* */
    int32_t mode;           /* potential walls: MAVI_WMODE_* */
    int32_t pot_kind;       /* potential walls: MAVI_POT_LJ or MAVI_POT_HARMTRUNC */
    int32_t nlines;         /* LinesCfg: segments in `lines` */
    double geom[4];         /* rectangle: bottom_left x, y, length, height;
                               circle: center x, y, radius */
    double pot[4];          /* potential walls: LJ sigma, epsilon; HarmTrunc
                               k_rep, k_atr, dist_eq, dist_max */
    const double *lines;    /* LinesCfg: nlines x (p1x, p1y, p2x, p2y) */
} mavi_wall;

/* Everything the step needs from System/SpaceCfg/DynamicCfg/IntCfg/Chunks. */
typedef struct {
    int64_t n;              /* length(state.pos): particle slots */
    double bottom_left[2];  /* RectangleCfg.bottom_left */
    double size[2];         /* RectangleCfg.length, .height */
    int32_t wall;           /* MAVI_WALL_* */
    int32_t potential;      /* MAVI_POT_* */
    double pot[4];          /* LJ: sigma, epsilon, rcut (<= 0: none, the
                               reference's LenJonesCfg); HT: k_rep, k_atr,
                               dist_eq, dist_max; Szabo: k_rep, k_adh, r_eq,
                               r_max; RTP: sigma, epsilon */
    double dt;              /* IntCfg.dt */
    int32_t cols, rows;     /* ChunksCfg(num_cols, num_rows); 0,0 = chunks nothing */
    int32_t nc;             /* chunk capacity (src/chunks.jl:32-34), checked */
    int32_t reserved0;
    double radius;          /* particle_radius(dynamic_cfg) for RigidWalls */
    /* Rings (potential == MAVI_POT_RINGS) */
    int32_t nmax;           /* num_max_particles(state) */
    int32_t nrings;         /* size(rings_pos, 2) */
    int32_t ntypes;         /* RingsCfg.num_types */
    int32_t reserved1;
    const int32_t *ring_types; /* state.types (1-based), NULL if untyped */
    const int32_t *type_np;    /* particles per ring for each type */
    const double *type_par;    /* per type: p0, relax_time, vo, mobility,
                                  rot_diff, k_area, k_spring, l_spring */
    const double *interaction; /* ntypes x ntypes HarmTrunc (k_rep, k_atr,
                                  dist_eq, dist_max), column-major like
                                  InteractionMatrix.matrix */
    uint64_t seed;          /* device noise stream (Philox4x32-10) */
    double active[4];       /* Szabo: vo, mobility, relax_time, rot_diff;
                               RTP: vo, tumble_rate (self-propelled only) */
    int32_t nwalls;         /* > 0: the walls list replaces `wall`; the
                               rectangle above is then the chunk space, the
                               bounding box of every geometry (get_chunks,
                               src/systems.jl:14-28) */
    int32_t reserved2;
    const mavi_wall *walls; /* read during mavi_sys_create only */
} mavi_sys_desc;

/* Library / context ------------------------------------------------------ */
int mavi_abi_version(void);
/* One context per host task; it owns one HIP stream on `device`. */
int mavi_ctx_create(int device, mavi_ctx **out);
int mavi_ctx_destroy(mavi_ctx *ctx);
const char *mavi_last_error(const mavi_ctx *ctx);

/* System: replaces System(...) construction, src/systems.jl:73-114 and
 * RingsSystem, src/rings/rings.jl:231-291 (device side only). */
int mavi_sys_create(mavi_ctx *ctx, const mavi_sys_desc *desc, mavi_sys **out);
int mavi_sys_destroy(mavi_sys *sys);

/* State transfer (AoS host <-> device).  vel is SecondLawState.vel, pol is
 * RingsState.pol (nrings angles) or SelfPropelledState.pol_angle (n angles);
 * either may be NULL. */
int mavi_sys_upload(mavi_sys *sys, const double *pos, const double *vel, const double *pol);
int mavi_sys_download(mavi_sys *sys, double *pos, double *vel, double *pol);

/* update_chunks!(chunks), src/chunks.jl:150-161 (+ src/integration.jl:54-59,
 * src/rings/integration.jl:18-23).  Bins the current positions; when the
 * arrays are given, exports chunks.chunk_particles (nc, rows, cols) and
 * chunks.num_particles_in_chunk (rows, cols), column-major, 1-based ids,
 * exactly the reference layout. */
int mavi_bin(mavi_sys *sys, int64_t *chunk_particles, int64_t *counts);

/* clean_forces! + update_chunks! + calc_forces!(system), src/integration.jl:112-226
 * (plain), forces!(system) src/rings/integration.jl:197-226 (rings: pair +
 * springs + area).  Writes get_forces(system) to forces (2n doubles). */
int mavi_forces(mavi_sys *sys, double *forces);

/* get_forces(system) as left by the last step (2n doubles). */
int mavi_get_forces(mavi_sys *sys, double *forces);

/* nsteps x newton_step! (src/integration.jl:507-515) for LJ/HT systems,
 * nsteps x Rings.Integration.step! (src/rings/integration.jl:522-543) for
 * rings, or nsteps x szabo_step! / rtp_step! (src/integration.jl:517-535)
 * for self-propelled systems, with device noise.  State stays resident. */
int mavi_step(mavi_sys *sys, int64_t nsteps);

/* nsteps x szabo_step! / rtp_step! (src/integration.jl:517-535) for
 * self-propelled systems (PeriodicWalls).  noise replaces the global RNG,
 * per step and particle id (0-based):
 *   Szabo: noise[step*n + i]          standard normal, the randn() of :459
 *   RTP:   noise[(step*n + i)*2 + 0]  uniform [0,1), the rand() of :484
 *          noise[(step*n + i)*2 + 1]  uniform [0,1), the rand() of :486
 * NULL draws from the device Philox4x32-10 stream (the same as mavi_step). */
int mavi_sp_step(mavi_sys *sys, int64_t nsteps, const double *noise);

/* Rings step with host-supplied standard-normal draws: noise[step*nrings + r]
 * replaces randn(system.rng) of src/rings/integration.jl:348.  NULL = zero. */
int mavi_rings_step(mavi_sys *sys, int64_t nsteps, const double *noise);

/* Contact neighbours of a rings system: ParticleNeighbors
 * (src/rings/neighbors.jl:47-137) as calc_interaction fills them during
 * forces! (src/rings/integration.jl:32-45).  With enable, every forces!
 * (each step, and mavi_forces) records for each particle the others closer
 * than tol * 2 * particle_radius of the pair's interaction; only_rings
 * keeps pairs of different rings only (type = :rings). */
int mavi_rings_neighbors_cfg(mavi_sys *sys, int enable, double tol, int only_rings);
/* The last forces!' neighbours: counts (nslots) and, when given, lists
 * (nslots x 15, per particle 15 entries of 1-based ids in ascending order,
 * -1 padded; the reference's list holds the same set in its traversal
 * order).  More than 15 neighbours fails the step with
 * MAVI_E_CELL_OVERFLOW, as the reference's BoundsError. */
int mavi_rings_neighbors_get(mavi_sys *sys, int64_t *counts, int64_t *lists);

/* Invasions of rings: check_invasions! (src/rings/integration.jl:379-509)
 * with the ring chunks r_chunks (src/rings/rings.jl:188-202: r_rows x r_cols
 * chunks of capacity r_nc over the chunk space, binning the rings' centres
 * of mass; r_cols = r_rows = 0: r_chunks = nothing, every pair of rings) and,
 * when scheduled, InvasionsCfg(steps_to_update) (src/rings/configs.jl:334-
 * 350): a step checks when num_steps - last_check >= steps_to_update
 * (update_invasions!, :511-520), last_check being the step count of this
 * call.  The check uses the ring points and centres of mass of the step's
 * forces! (RingsInfo.continuos_pos and cms).  Clears the list; with
 * steps_to_update <= 0 checks at once, as the constructor does.  A centre of
 * mass off the ring grid fails with MAVI_E_OUT_OF_BOX, a full ring chunk
 * with MAVI_E_CELL_OVERFLOW (the reference's BoundsError). */
int mavi_rings_invasions_cfg(mavi_sys *sys, int scheduled, int64_t steps_to_update, int32_t r_cols,
                             int32_t r_rows, int32_t r_nc);
/* check_invasions!(system) now, at the last step's ring points and centres
 * of mass (after creation: the initial ones). */
int mavi_rings_invasions_check(mavi_sys *sys);
/* InvasionsInfo.list of the last check (src/rings/rings.jl:86-100): *count
 * records and up to cap of them as (invasor, invaded, p_id) triples of
 * 1-based ring ids and the invading particle's scalar index, ordered by
 * (invasor, invaded, p_id); the reference's list holds the same set in the
 * order of its chunk traversal. */
int mavi_rings_invasions_get(mavi_sys *sys, int64_t *count, int64_t *records, int64_t cap);

/* Variable ring population (VarRingsIds, src/rings/states.jl:20-39,
 * 173-223) with sources and sinks (src/rings/sources.jl), processed in list
 * order at the start of every step (update_sources!,
 * src/rings/integration.jl:353-358). */
#define MAVI_SRC_SOURCE 0
#define MAVI_SRC_SINK 1
#define MAVI_GEOM_RECT 0   /* g = bottom_left x, y, length, height */
#define MAVI_GEOM_CIRCLE 1 /* g = centre x, y, radius */
typedef struct {
    int32_t kind;             /* MAVI_SRC_SOURCE / MAVI_SRC_SINK */
    int32_t geom;             /* sink: SinkCfg.geometry_cfg, MAVI_GEOM_* */
    double g[4];
    int32_t nbox;             /* source: length(Source.bbox) (at most 64 over all sources) */
    int32_t random_pol;       /* 1: spawn_pol = :random (2 pi u, u from mavi_rings_pol_draws) */
    double pad;               /* SourceCfg.pad */
    double pol;               /* a fixed spawn_pol */
    const double *bbox;       /* nbox x 4: Source.bbox (bottom_left x, y, length, height) */
    const double *spawn;      /* nbox x nmax x 2: Source.bbox_spawn_pos (add_ring! writes every slot) */
    const int32_t *chunk_off; /* nbox + 1 offsets into chunk_ids (ChunksChecker.ids), NULL: PosChecker */
    const int32_t *chunk_ids; /* chunk (row, col) as (row - 1) * cols + (col - 1) */
} mavi_ring_source;

/* Make a rings system's population variable: the ActiveState mask
 * (nrings, NULL: every ring) and its sources and sinks (at most 32 sinks).
 * Call once, after the first upload: the constructor's update_ids!,
 * centres of mass (0 for inactive rings) and first update_chunks! follow.
 * Steps then run update_cms!, the sources and sinks, update_ids! and the
 * rest of step! over the active rings; every step reads back a few words
 * to decide (no graph replays).  The host noise of mavi_rings_step stays
 * indexed by ring slot. */
int mavi_rings_var_cfg(mavi_sys *sys, const int32_t *mask, int32_t nsrc, const mavi_ring_source *src);
/* Uniform draws u in [0, 1) for random spawn polarisations (2 pi u),
 * consumed in spawn order (one per empty box, found slot or not, as
 * get_spawn_pol runs before add_ring!); appended to the pool.  Once the
 * pool is empty the draws come from a built-in stream (Philox4x32-10 keyed
 * by the system's seed and the stream position). */
int mavi_rings_pol_draws(mavi_sys *sys, const double *u, int64_t n);
/* Position of the built-in spawn polarisation stream (to resume a saved run). */
int mavi_rings_pol_stream_set(mavi_sys *sys, int64_t position);
/* VarRingsIds: mask (nrings), ids (nrings, 1-based; past num_active the
 * entries an earlier update left), uids (nrings), num_active, the pool draws
 * used so far and the built-in stream's position; any may be NULL. */
int mavi_rings_var_get(mavi_sys *sys, int32_t *mask, int64_t *ids, int64_t *uids, int64_t *num_active,
                       int64_t *pol_used, int64_t *pol_stream);

/* RingsInfo: areas (nrings), cms (2*nrings), continuos_pos (2*nmax*nrings),
 * src/rings/rings.jl:118-128; any may be NULL. */
int mavi_rings_info(mavi_sys *sys, double *areas, double *cms, double *continuous_pos);

/* kinetic_energy(state) and potential_energy(system, ::LenJonesCfg),
 * src/quantities.jl:12-18, 46-66 (PE over all pairs, as the reference). */
int mavi_observables(mavi_sys *sys, double *ke, double *pe);

/* nsteps x newton_step!, recording kinetic and potential energy after every
 * step (ke, pe: nsteps doubles each; pe may be NULL). */
int mavi_run_record(mavi_sys *sys, int64_t nsteps, double *ke, double *pe);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * steps launch kernels directly (no graph) and every launch of the timed
 * kernel classes is bracketed by events.  mavi_prof_get returns, for kernel
 * class `which`, the summed duration in ms and the number of launches. */
#define MAVI_PROF_BIN 0
#define MAVI_PROF_FORCE 1    /* pair-force kernels (full evaluations) */
#define MAVI_PROF_RINGS 2    /* per-ring springs/area/overdamped kernel */
#define MAVI_PROF_FORCE_REUSE 3  /* first evaluations of Verlet steps that may reuse
                                    the previous step's second one (MAVI_OPT_REUSE_FORCES) */
#define MAVI_PROF_NCLASS 4
int mavi_prof_enable(mavi_sys *sys, int on);
int mavi_prof_get(mavi_sys *sys, int which, double *ms_total, int64_t *launches);

/* Force-kernel tile shape: 0 = library default; otherwise the workgroup size
 * (64, 128, 256, 512, 1024) used by the pair-force kernels.  Used by the
 * autotuner. */
int mavi_sys_set_tile(mavi_sys *sys, int32_t workgroup);

/* Execution options (no effect on results beyond floating-point summation
 * order; none changes the reference semantics):
 *   MAVI_OPT_TWO_PHASE  1: finite-range interactions first keep the in-range
 *                       candidates, then evaluate forces on them (default 0)
 *   MAVI_OPT_GRAPHS     1 (default): multi-step calls replay a captured
 *                       two-step hipGraph
 *   MAVI_OPT_LDS_TILES  1: pair kernels stage chunk-row segments in LDS;
 *                       0 (default): per-particle gathers from global memory
 *   MAVI_OPT_REUSE_FORCES  1 (default): the first force evaluation of a
 *                       newton_step! takes the previous step's second
 *                       evaluation for every wavefront whose particles' 3x3
 *                       chunk blocks kept their members (same positions,
 *                       same candidates in the same order: bit-identical);
 *                       0: always evaluate */
/* Position of the device noise stream (Philox4x32-10 keyed by the seed and
 * a step counter that every step advances): saving both with a state and
 * setting them on the system rebuilt from it continues the stream exactly
 * (the role the serialized system.rng plays in save_system/load_system,
 * src/serder.jl:62-66, 140-146). */
int mavi_sys_noise_get(mavi_sys *sys, uint64_t *seed, uint64_t *counter);
int mavi_sys_noise_set(mavi_sys *sys, uint64_t seed, uint64_t counter);

#define MAVI_OPT_TWO_PHASE 0
#define MAVI_OPT_GRAPHS 1
#define MAVI_OPT_LDS_TILES 2
#define MAVI_OPT_REUSE_FORCES 3
/*   MAVI_OPT_NLIST      1 (default): newton_step! of a finite-range plain
 *                       system (LJ with a cutoff, HarmTrunc) whose chunks are
 *                       at least the range wide runs on Verlet neighbour
 *                       lists: every pair within range, the step's chunks
 *                       kept for the second evaluation (pairs whose
 *                       step-start chunks are not adjacent left out), forces
 *                       summed in ascending partner id; lists rebuilt when a
 *                       particle has moved half the skin.  0: the chunk gather
 *   MAVI_OPT_NLIST_SKIN skin in thousandths of the range (0: 150)
 *   MAVI_OPT_NLIST_BLOCK steps between displacement checks (0: as many as
 *                       the displacement bound allows at the last measured
 *                       step length, at most 60) */
#define MAVI_OPT_NLIST 4
#define MAVI_OPT_NLIST_SKIN 5
#define MAVI_OPT_NLIST_BLOCK 6
int mavi_sys_set_option(mavi_sys *sys, int32_t option, int32_t value);

/* Neighbour lists: info[0] the system has them (eligible), [1] they run,
 * [2] rebuilds and [3] rolled-back blocks so far. */
int mavi_nlist_info(mavi_sys *sys, int64_t *info);

/* Autotuner: times `trial_steps` steps of the system's own dynamics for every
 * pair-kernel shape (workgroup 64..512 x LDS tiles x two-phase list), keeps
 * the fastest and restores the state it started from.  best_us (nullable)
 * receives the chosen shape's time per step; shape (nullable, 3 ints)
 * receives {workgroup, lds_tiles, two_phase}. */
int mavi_autotune(mavi_sys *sys, int32_t trial_steps, double *best_us, int32_t *shape);

/* Replica batches ---------------------------------------------------------
 * Many independent systems of uneven size stepped together: the device side
 * of run_experiment_batch (src/experiments.jl:246-300), where every replica
 * is its own System stepped by its own step function.  Every phase of a step
 * is one launch over all replicas of a kind, and multi-step runs replay a
 * captured graph of `graph_steps` steps.  Each replica's results are
 * bit-identical to the same descriptor stepped alone with mavi_step.
 * Replicas need chunks (cols, rows > 0); rings replicas need nmax <= 32. */
typedef struct mavi_batch mavi_batch;
int mavi_batch_create(mavi_ctx *ctx, mavi_batch **out);
int mavi_batch_destroy(mavi_batch *batch);
/* Append a replica (descriptor as for mavi_sys_create; host tables are read
 * during the call only).  *index receives its 0-based index. */
int mavi_batch_add_replica(mavi_batch *batch, const mavi_sys_desc *desc, int32_t *index);
/* Allocate the device arrays; no replicas may be added afterwards. */
int mavi_batch_finalize(mavi_batch *batch);
/* State of one replica, as mavi_sys_upload / mavi_sys_download. */
int mavi_batch_upload(mavi_batch *batch, int32_t index, const double *pos, const double *vel, const double *pol);
int mavi_batch_download(mavi_batch *batch, int32_t index, double *pos, double *vel, double *pol);
int mavi_batch_get_forces(mavi_batch *batch, int32_t index, double *forces);
/* nsteps of every replica's own dynamics (newton_step! or Rings step! with
 * the device rotational noise).  A failing replica is named in
 * mavi_last_error. */
int mavi_batch_run(mavi_batch *batch, int64_t nsteps);
/* The batch's step counter (shared by its replicas' noise streams, each
 * keyed by its own seed): replicas resumed together at step k continue with
 * the counter set to k. */
int mavi_batch_noise_get_counter(mavi_batch *b, uint64_t *counter);
int mavi_batch_noise_set_counter(mavi_batch *b, uint64_t counter);
/* Batch execution options:
 *   MAVI_BOPT_WG_PLAIN / MAVI_BOPT_WG_RINGS  pair-kernel workgroup (64..1024)
 *       for LJ/HarmTrunc replicas / rings replicas (default 256)
 *   MAVI_BOPT_LIST_PLAIN / MAVI_BOPT_LIST_RINGS  two-phase candidate list
 *       (as MAVI_OPT_TWO_PHASE, default 0)
 *   MAVI_BOPT_GRAPH_STEPS  steps per captured graph (even, >= 2; default 100)
 *   MAVI_BOPT_GRAPHS       1 (default): replay graphs; 0: direct launches */
#define MAVI_BOPT_WG_PLAIN 0
#define MAVI_BOPT_WG_RINGS 1
#define MAVI_BOPT_LIST_PLAIN 2
#define MAVI_BOPT_LIST_RINGS 3
#define MAVI_BOPT_GRAPH_STEPS 4
#define MAVI_BOPT_GRAPHS 5
int mavi_batch_set_option(mavi_batch *batch, int32_t option, int32_t value);
/* Sizes: info[0] replicas, [1] particle slots (sum of n), [2] active
 * particles, [3] padded sorted slots, [4] chunks incl. one pad chunk per
 * replica, [5] rings. */
int mavi_batch_info(mavi_batch *batch, int64_t *info);
/* Kernel timing as mavi_prof_enable / mavi_prof_get (classes MAVI_PROF_*). */
int mavi_batch_prof_enable(mavi_batch *batch, int on);
int mavi_batch_prof_get(mavi_batch *batch, int which, double *ms_total, int64_t *launches);
/* Autotuner: for LJ/HarmTrunc replicas and then rings replicas, times
 * trial_steps steps for every workgroup 64..512 x list on/off, keeps the
 * fastest and restores the state.  shape (nullable, 4 ints) receives
 * {wg_plain, list_plain, wg_rings, list_rings}. */
int mavi_batch_autotune(mavi_batch *batch, int32_t trial_steps, double *best_us, int32_t *shape);

/* Learned propagator: network --------------------------------------------
 * No counterpart in the reference (an addition beside the plain path).  An
 * MLP MAVI_MLP_IN -> 128 -> 128 -> 128 -> MAVI_MLP_OUT with GELU; BF16
 * weights and activations, FP32 accumulation on the matrix cores.  Features
 * and weights cross the boundary as BF16 bit patterns (uint16_t); parameters
 * as the FP32 master copy, layout W0^T[128][160], b0[128], W1^T[128][128],
 * b1, W2^T[128][128], b2, W3^T[2][128], b3[2] (W^T: out x in, row-major). */
#define MAVI_MLP_IN 160
#define MAVI_MLP_HIDDEN 128
#define MAVI_MLP_OUT 2
#define MAVI_MLP_NPARAMS 53890
typedef struct mavi_mlp mavi_mlp;
int mavi_mlp_create(mavi_ctx *ctx, uint64_t seed, mavi_mlp **out);
int mavi_mlp_destroy(mavi_mlp *mlp);
int mavi_mlp_num_params(void);
int mavi_mlp_get_params(mavi_mlp *mlp, float *params);
int mavi_mlp_set_params(mavi_mlp *mlp, const float *params);
/* y[m][2] = MLP(x[m][160]); fused = 1: the inference kernel (weights in
 * LDS), 0: the training forward (GEMM chain). */
int mavi_mlp_forward(mavi_mlp *mlp, int64_t m, const uint16_t *x, float *y, int32_t fused);
/* mavi_mlp_forward (fused inference kernel) that allocates nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (inference in tight serving loops, code that must not call
 * hipMalloc).  x and y are staged through a fixed device array of the
 * library, one per GPU, sized for MAVI_MLP_NOALLOC_ROWS rows; more rows fail
 * with MAVI_E_BADARG.  Calls on one GPU share that array: calls that overlap
 * on one GPU (two contexts, two host threads) race. */
#define MAVI_MLP_NOALLOC_ROWS 8192
int mavi_mlp_forward_noalloc(mavi_mlp *mlp, int64_t m, const uint16_t *x, float *y);
/* Loss L = 1/(2m) sum (y - t)^2 and its gradient w.r.t. every parameter
 * (MAVI_MLP_NPARAMS floats, parameter layout); either output may be NULL. */
int mavi_mlp_loss_grad(mavi_mlp *mlp, int64_t m, const uint16_t *x, const float *t, double *loss, float *grad);
/* Device-resident training set: rows of features and targets. */
int mavi_mlp_dataset_clear(mavi_mlp *mlp);
int mavi_mlp_dataset_append(mavi_mlp *mlp, int64_t n, const uint16_t *x, const float *t);
int mavi_mlp_dataset_size(mavi_mlp *mlp, int64_t *n);
int mavi_mlp_dataset_get(mavi_mlp *mlp, int64_t first, int64_t n, uint16_t *x, float *t);
/* steps x (minibatch of `batch` rows drawn with replacement from the data
 * set by (seed, step), forward + backward, Adam with learning rate lr);
 * losses (nullable, steps floats) receives each step's minibatch loss. */
int mavi_mlp_train(mavi_mlp *mlp, int64_t steps, int32_t batch, float lr, uint64_t seed, float *losses);
/* Test hook for the matrix-core tiles: c[m][n] = sum_k A(m,k) B(k,n) in FP32
 * with A = a[m][k] (ta = 0) or a[k][m] (ta = 1), B = b[k][n] (tb = 0) or
 * b[n][k] (tb = 1), all BF16 bit patterns. */
int mavi_mlp_gemm_check(mavi_ctx *ctx, int32_t m, int32_t n, int32_t k, int32_t ta, int32_t tb, const uint16_t *a,
                        const uint16_t *b, float *c);
/* Average time of the inference kernel over m device-resident rows. */
int mavi_mlp_time_forward(mavi_mlp *mlp, int64_t m, int32_t reps, double *us_per_call);

/* Learned propagator: stepping ----------------------------------------------
 * No counterpart in the reference.  Binds a rings system (periodic, chunks
 * >= r_feat wide, full rings with nmax <= 32) and a network.  One learned
 * step stands for k Rings step! calls: features of every particle in its
 * ring's polarisation frame (32 self features, 32 nearest neighbours within
 * r_feat), displacement = k dt (v0 p + mu F) + target_scale * R^T MLP(x),
 * gated by displacement, overlap and ring-energy checks; a rejected step is
 * replaced by k plain FP64 steps.  Feature rows are in particle-id order. */
typedef struct {
    int32_t k;             /* plain steps one learned step stands for */
    int32_t reserved;
    double r_feat;         /* neighbour radius of the features */
    double target_scale;   /* s_t: network output unit of displacement */
    double d_min;          /* overlap gate distance (<= 0: 0.3 x the smallest dist_eq) */
    double max_disp;       /* displacement gate (<= 0: (r_feat - d_min)/2, the largest allowed) */
    double e_tol;          /* per-ring spring + area energy rise gate (<= 0: 1.0) */
    uint64_t seed;         /* rotational-noise stream of accepted learned steps */
} mavi_learned_desc;
typedef struct mavi_learned mavi_learned;
/* mlp may be NULL: only the network-free mode of mavi_learned_run is then available. */
int mavi_learned_create(mavi_sys *sys, mavi_mlp *mlp, const mavi_learned_desc *desc, mavi_learned **out);
int mavi_learned_destroy(mavi_learned *learned);
/* Bin, forces!, features at the current state: x (n x 160 BF16, nullable) and
 * the explicit k dt displacement (n x 2, nullable); the staging copies read
 * by mavi_sys_download / mavi_get_forces / mavi_rings_info match them. */
int mavi_learned_features(mavi_learned *learned, uint16_t *x, double *euler);
/* Training data from the plain method: nsnap x (features, k plain steps with
 * the device noise, targets), rows_per_snap rows per snapshot (>= n: every
 * particle in id order, else drawn by (seed, snapshot, row)) appended to the
 * network's data set, then `stride` further plain steps. */
int mavi_learned_collect(mavi_learned *learned, int32_t nsnap, int32_t stride, int32_t rows_per_snap, uint64_t seed);
/* nsteps learned steps (each k plain-step equivalents); use_net = 0 runs the
 * same gated step without the network correction.  *accepted (nullable)
 * receives how many passed the gate. */
int mavi_learned_run(mavi_learned *learned, int64_t nsteps, int32_t use_net, int64_t *accepted);
/* Totals since creation (7 values): accepted, rejected, then rejections by
 * gate check: displacement, overlap, ring energy, more than 32 neighbours,
 * non-finite network output (a step may fail several). */
int mavi_learned_stats(mavi_learned *learned, int64_t *stats);

#ifdef __cplusplus
}
#endif
#endif
