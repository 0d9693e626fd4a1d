/*
 * mavi_oracle.c -- CPU restatement of Mavi.jl's step hot path (TEST
 * INFRASTRUCTURE ONLY: the checker for libmavi_hip.so, never the product).
 *
 * Compiled with -ffp-contract=off: Julia does not contract a*b+c into an FMA
 * unless @fastmath/@muladd is used, so every expression here is evaluated as
 * the Julia source writes it, left to right, one rounding per operation.
 *
 * How each part is pinned (tests/test_oracle_*.py), independently of the
 * formulas restated here:
 *   or_jl_div / or_cell        exact rational arithmetic (fractions.Fraction)
 *                              on adversarial inputs: exact multiples of the
 *                              cell size, +-1 ulp, the top/right edge clamp.
 *   or_chunk_capacity          the capacities the reference's own configs give
 *                              (32 for 17x17 over 1024 LJ particles etc.).
 *   or_bin                     invariants: every active id exactly once,
 *                              ascending inside a cell, counts sum to N, each
 *                              id in the cell recomputed exactly.
 *   or_neighbors               every unordered pair of 8-adjacent cells
 *                              appears exactly once (grids >= 3x3), brute
 *                              force enumeration.
 *   or_pair_force (LJ, HT)     closed forms: LJ zero at 2^(1/6) sigma and
 *                              24 eps/sigma at sigma, minus the derivative of
 *                              the LJ energy (finite differences); HT
 *                              piecewise values; min-image at +-L/2.
 *   or_calc_forces             brute-force all pairs == cell list for finite
 *                              range potentials (the reference's own
 *                              check_chunks technique), Newton's third law
 *                              (sum F = 0), Threaded == Sequencial.
 *   or_newton_steps            free flight x + v dt exactly, two-body harmonic
 *                              oscillation against the analytic solution,
 *                              energy conservation O(dt^2), time reversal,
 *                              wall reflection / wrap closed forms.
 *   or_kinetic/or_potential    brute force on tiny inputs in Python floats.
 *   Szabo force                minus the gradient of U(d) = f_mod (d^3/3 - r_eq
 *                              d^2/2) by finite differences on each branch, zero
 *                              at r_eq and beyond r_max, Newton's third law.
 *   or_sp_steps                free self-propulsion x + k vo pol dt, the
 *                              polarisation turning towards the force,
 *                              var(dtheta) = 2 Dr dt, tumbles at u < rate dt
 *                              to 2 pi v exactly, periodic wrap.
 *   or_rings_invasions         point_line_intersect: the parametric crossing
 *                              of the +x ray with the open segment on random
 *                              segments; hand polygons (square, concave L,
 *                              a point level with a vertex, horizontal and
 *                              vertical edges); ring pairs: a winding-number
 *                              point-in-polygon over every chunk-adjacent
 *                              pair (random overlapping rings), chunks =
 *                              all pairs where the grid is coarse.
 *   or_rings_neighbors         brute force over all pairs on random ring
 *                              gases (same sets), counts = list lengths,
 *                              symmetry (j in list(i) <=> i in list(j)).
 *   walls list                 rigid circle: speed kept, normal component
 *                              reversed (centred at the origin); slippery
 *                              circle / line / corner: ends at distance r;
 *                              potential walls: minus the finite-difference
 *                              gradient of the wall energy; list order.
 *   rings                      regular n-gon area n l^2/(4 tan(pi/n)), spring
 *                              force zero at rest length, area force zero at
 *                              A0 and equal to -dE/dx by finite differences,
 *                              unwrap continuity, overdamped drift closed form
 *                              with rot_diff = 0.
 */
#include "mavi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Julia Base arithmetic used by the reference                              */
/* ------------------------------------------------------------------------ */

/* div(a, b) for Float64 in Julia Base (RoundToZero): the remainder is taken
 * exactly with rem (= C fmod), the rest divided and rounded to nearest.
 * Used by update_particle_chunk!, src/chunks.jl:129-130. */
double or_jl_div(double a, double b) { return nearbyint((a - fmod(a, b)) / b); }

/* src/chunks.jl:32-34:
 *   nc = (ceil(0.5*chunk_length/particle_r) + 1) * ((ceil(0.5*chunk_height/particle_r) + 1))
 *   nc = trunc(Int, ceil(nc*2)) */
int64_t or_chunk_capacity(double cl, double ch, double r) {
    double nc = (ceil(0.5 * cl / r) + 1.0) * (ceil(0.5 * ch / r) + 1.0);
    return (int64_t)trunc(ceil(nc * 2.0));
}

/* update_particle_chunk!, src/chunks.jl:120-142, on a grid of rows x cols
 * chunks of size cl x ch over the system's chunk space. */
static int cell_in(const or_sys *s, int64_t rows, int64_t cols, double cl, double ch, double x, double y,
                   int64_t *row, int64_t *col) {
    double H = s->size[1];
    double qr = or_jl_div((-y + s->bl[1]) + H, ch);
    double qc = or_jl_div(x - s->bl[0], cl);
    if (!isfinite(qr) || !isfinite(qc)) return OR_E_OUT_OF_BOX; /* trunc(Int, NaN) throws */
    int64_t r = (int64_t)qr + 1;
    int64_t c = (int64_t)qc + 1;
    r -= (r == rows + 1) ? 1 : 0;
    c -= (c == cols + 1) ? 1 : 0;
    *row = r;
    *col = c;
    if (r < 1 || r > rows || c < 1 || c > cols) return OR_E_OUT_OF_BOX; /* BoundsError */
    return OR_OK;
}

int or_cell(const or_sys *s, double x, double y, int64_t *row, int64_t *col) {
    return cell_in(s, s->rows, s->cols, s->cl, s->ch, x, y, row, col);
}

/* get_particles_ids: all slots for plain states (src/states.jl:213-228 with
 * active_state = nothing), FixRingsIds.p_ids for rings (src/rings/states.jl:46-62):
 * untyped rings use eachindex(pos); typed rings list the first
 * num_particles[type] slots of every ring. */
static int32_t ring_np(const or_sys *s, int32_t ring /*1-based*/) {
    if (s->types == NULL) return s->np_type[0];
    return s->np_type[s->types[ring - 1] - 1];
}

static inline int ring_active(const or_sys *s, int32_t ring /*1-based*/) {
    return s->var == NULL || s->var->mask[ring - 1] != 0;
}

int64_t or_particle_ids(const or_sys *s, int64_t *ids) {
    if (s->kind != OR_KIND_RINGS || (s->types == NULL && s->var == NULL)) {
        if (ids)
            for (int64_t i = 0; i < s->n; i++) ids[i] = i + 1;
        return s->n;
    }
    /* typed rings: FixRingsIds.p_ids; a variable population: the p_ids of
     * calc_active_ids! (src/rings/states.jl:200-223), active rings ascending */
    int64_t k = 0;
    for (int32_t r = 1; r <= s->nrings; r++) {
        if (!ring_active(s, r)) continue;
        int64_t base = (int64_t)(r - 1) * s->nmax + 1; /* to_scalar_idx(r, 1), src/rings/states.jl:137-139 */
        for (int32_t i = 1; i <= ring_np(s, r); i++) {
            if (ids) ids[k] = base + i - 1;
            k++;
        }
    }
    return k;
}

/* update_chunks!: src/chunks.jl:150-161 (+ :144-146 append), visiting
 * get_particles_ids in ascending order (src/integration.jl:54-59,
 * src/rings/integration.jl:18-23). */
int or_bin(const or_sys *s, const double *pos, int64_t *chunk_particles, int64_t *counts) {
    const int64_t rows = s->rows, cols = s->cols, nc = s->nc;
    memset(counts, 0, sizeof(int64_t) * rows * cols);
    int64_t nids = or_particle_ids(s, NULL);
    int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (nids > 0 ? nids : 1));
    or_particle_ids(s, ids);
    int err = OR_OK;
    for (int64_t k = 0; k < nids; k++) {
        int64_t i = ids[k];
        int64_t r, c;
        int e = or_cell(s, pos[2 * (i - 1)], pos[2 * (i - 1) + 1], &r, &c);
        if (e) { err = e; break; }
        int64_t cell = (r - 1) + rows * (c - 1);
        int64_t p = counts[cell] + 1;
        if (p > nc) { err = OR_E_CELL_OVERFLOW; break; } /* BoundsError on chunk_particles */
        chunk_particles[(p - 1) + nc * cell] = i;
        counts[cell] += 1;
    }
    free(ids);
    return err;
}

/* get_neighbors: periodic src/chunks.jl:61-87, bounded src/chunks.jl:89-118. */
static int64_t wrap_id(int64_t x, int64_t n) { /* get_id, src/chunks.jl:64-73 */
    if (x == 0) return n;
    if (x % (n + 1) == 0) return 1;
    return x;
}

void or_neighbors(int32_t rows, int32_t cols, int32_t periodic, int64_t *nb, int32_t *nnb) {
#define SETNB(r_, c_, e_, R_, C_)                                  \
    do {                                                           \
        int64_t k_ = (int64_t)((c_)-1) * rows + ((r_)-1);          \
        nb[(k_ * 4 + (e_)) * 2 + 0] = (R_);                        \
        nb[(k_ * 4 + (e_)) * 2 + 1] = (C_);                        \
    } while (0)
    if (periodic) {
        for (int64_t i = 1; i <= rows; i++)
            for (int64_t j = 1; j <= cols; j++) {
                SETNB(i, j, 0, wrap_id(i + 1, rows), wrap_id(j, cols));
                SETNB(i, j, 1, wrap_id(i + 1, rows), wrap_id(j + 1, cols));
                SETNB(i, j, 2, wrap_id(i, rows), wrap_id(j + 1, cols));
                SETNB(i, j, 3, wrap_id(i - 1, rows), wrap_id(j + 1, cols));
                nnb[(j - 1) * rows + (i - 1)] = 4;
            }
        return;
    }
    for (int64_t i = 1; i <= rows - 1; i++) {
        SETNB(i, 1, 0, i + 1, 1);
        SETNB(i, 1, 1, i + 1, 2);
        SETNB(i, 1, 2, i, 2);
        nnb[(1 - 1) * rows + (i - 1)] = 3;
        for (int64_t j = 2; j <= cols - 1; j++) {
            SETNB(i, j, 0, i + 1, j - 1);
            SETNB(i, j, 1, i + 1, j);
            SETNB(i, j, 2, i + 1, j + 1);
            SETNB(i, j, 3, i, j + 1);
            nnb[(j - 1) * rows + (i - 1)] = 4;
        }
        SETNB(i, cols, 0, i + 1, cols - 1);
        SETNB(i, cols, 1, i + 1, cols);
        nnb[(cols - 1) * rows + (i - 1)] = 2;
    }
    for (int64_t j = 1; j <= cols - 1; j++) {
        SETNB(rows, j, 0, rows, j + 1);
        nnb[(j - 1) * rows + (rows - 1)] = 1;
    }
    nnb[(cols - 1) * rows + (rows - 1)] = 0;
#undef SETNB
}

/* calc_diff, src/integration.jl:37-48.  Periodic rectangle:
 *   dr = dr - (abs.(dr) .> (size_vec / 2)) .* copysign.(size_vec, dr) */
void or_calc_diff(const or_sys *s, const double *r1, const double *r2, double *out) {
    for (int d = 0; d < 2; d++) {
        double dr = r1[d] - r2[d];
        if (s->periodic) {
            double half = s->size[d] / 2.0;
            double mask = (fabs(dr) > half) ? 1.0 : 0.0;
            dr = dr - mask * copysign(s->size[d], dr);
        }
        out[d] = dr;
    }
}

/* ------------------------------------------------------------------------ */
/* Pair forces                                                              */
/* ------------------------------------------------------------------------ */

/* potential_force(dr, dist, ::LenJonesCfg), src/configs.jl:389-397:
 *   fmod = 4*epsilon*(12*sigma^12/dist^13 - 6*sigma^6/dist^7); fmod/dist * dr
 * The optional cutoff (rcut > 0) returns zero(dr) for dist > rcut, the form
 * the reference uses for its truncated LJ (RunTumbleCfg, src/integration.jl:96-103). */
static void lj_force(const double *p, const double *dr, double *f) {
    double sigma = p[0], eps = p[1], rcut = p[2];
    double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
    if (rcut > 0.0 && dist > rcut) { f[0] = 0.0; f[1] = 0.0; return; }
    double fmod = 4.0 * eps * (12.0 * pow(sigma, 12.0) / pow(dist, 13.0) - 6.0 * pow(sigma, 6.0) / pow(dist, 7.0));
    double q = fmod / dist;
    f[0] = q * dr[0];
    f[1] = q * dr[1];
}

/* potential_force(dr, dist, ::HarmTruncCfg), src/configs.jl:354-368. */
static void ht_force(const double *p, const double *dr, double *f) {
    double k_rep = p[0], k_atr = p[1], dist_eq = p[2], dist_max = p[3];
    double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
    if (dist > dist_max) { f[0] = 0.0; f[1] = 0.0; return; }
    double fmod;
    if (dist < dist_eq) fmod = -k_rep * (dist / dist_eq - 1.0);
    else fmod = -k_atr * (dist / dist_eq - 1.0);
    double q = fmod / dist;
    f[0] = q * dr[0];
    f[1] = q * dr[1];
}

/* calc_interaction(i, j, ::SzaboCfg), src/integration.jl:68-86:
 *   dist > r_max -> 0; f_mod = k_adh/r_eq (dist > r_eq) or k_rep/(r_max - r_eq);
 *   -f_mod * (dist - r_eq) * dr, evaluated ((-f_mod) * r) * dr. */
static void szabo_force(const double *p, const double *dr, double *f) {
    double k_rep = p[0], k_adh = p[1], r_eq = p[2], r_max = p[3];
    double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
    if (dist > r_max) { f[0] = 0.0; f[1] = 0.0; return; }
    double f_mod = (dist > r_eq) ? k_adh / r_eq : k_rep / (r_max - r_eq);
    double r = dist - r_eq;
    double q = -f_mod * r;
    f[0] = q * dr[0];
    f[1] = q * dr[1];
}

static inline int32_t ring_of(const or_sys *s, int64_t idx) { /* get_ring_id, src/rings/states.jl:142 */
    return (int32_t)((idx - 1) / s->nmax + 1);
}

static const double *inter_of(const or_sys *s, int32_t r1, int32_t r2) {
    /* get_interaction_cfg, src/rings/configs.jl:79-87 */
    if (s->types == NULL) return s->inter;
    int32_t t1 = s->types[r1 - 1], t2 = s->types[r2 - 1];
    return s->inter + 4 * ((t1 - 1) + (int64_t)s->ntypes * (t2 - 1));
}

/* Rings calc_interaction + calc_interaction_force, src/rings/integration.jl:32-77. */
static void rings_force(const or_sys *s, const double *pos, int64_t i, int64_t j, double *f) {
    int32_t ri = ring_of(s, i), rj = ring_of(s, j);
    const double *ic = inter_of(s, ri, rj);
    double dr[2];
    or_calc_diff(s, pos + 2 * (i - 1), pos + 2 * (j - 1), dr);
    double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
    f[0] = 0.0;
    f[1] = 0.0;
    if (dist > ic[3]) return;
    if (ri == rj) {
        int64_t diff = i > j ? i - j : j - i;
        int64_t np = ring_np(s, ri);
        if (diff == 1 || diff == np - 1) return;
    }
    double dist_eq = ic[2];
    double fmod;
    if (dist < dist_eq) {
        fmod = -ic[0] * (dist / dist_eq - 1.0);
    } else {
        if (ri == rj) fmod = 0.0;
        else fmod = -ic[1] * (dist / dist_eq - 1.0);
    }
    double q = fmod / dist;
    f[0] = q * dr[0];
    f[1] = q * dr[1];
}

void or_pair_force(const or_sys *s, const double *pos, int64_t i, int64_t j, double *f) {
    if (s->kind == OR_KIND_RINGS) { rings_force(s, pos, i, j, f); return; }
    double dr[2];
    or_calc_diff(s, pos + 2 * (i - 1), pos + 2 * (j - 1), dr);
    if (s->kind == OR_KIND_LJ || s->kind == OR_KIND_RTP) lj_force(s->p, dr, f);
    else if (s->kind == OR_KIND_SZABO) szabo_force(s->p, dr, f);
    else ht_force(s->p, dr, f);
}

/* ------------------------------------------------------------------------ */
/* calc_forces!                                                             */
/* ------------------------------------------------------------------------ */

static inline void acc_pair(const or_sys *s, const double *pos, int64_t p1, int64_t p2,
                            double *forces, double *absum) {
    double f[2];
    or_pair_force(s, pos, p1, p2, f);
    /* forces[p1_id] += f1; forces[p2_id] -= f1 (src/integration.jl:131-132) */
    forces[2 * (p1 - 1)] += f[0];
    forces[2 * (p1 - 1) + 1] += f[1];
    forces[2 * (p2 - 1)] -= f[0];
    forces[2 * (p2 - 1) + 1] -= f[1];
    if (absum) {
        double a = fabs(f[0]) + fabs(f[1]);
        absum[p1 - 1] += a;
        absum[p2 - 1] += a;
    }
}

/* One column of the Sequencial loop body, src/integration.jl:116-156. */
static void force_column(const or_sys *s, const double *pos, const int64_t *cp, const int64_t *counts,
                         const int64_t *nb, const int32_t *nnb, int64_t col, double *forces, double *absum) {
    const int64_t rows = s->rows, nc = s->nc;
    for (int64_t row = 1; row <= rows; row++) {
        int64_t k = (col - 1) * rows + (row - 1);
        int64_t np = counts[k];
        const int64_t *chunk = cp + nc * k;
        for (int64_t i = 0; i < np; i++) {
            int64_t p1 = chunk[i];
            for (int64_t j = i + 1; j < np; j++) acc_pair(s, pos, p1, chunk[j], forces, absum);
            for (int e = 0; e < nnb[k]; e++) {
                int64_t nr = nb[(k * 4 + e) * 2], ncol = nb[(k * 4 + e) * 2 + 1];
                int64_t nk = (ncol - 1) * rows + (nr - 1);
                int64_t nei_np = counts[nk];
                const int64_t *nei = cp + nc * nk;
                for (int64_t j = 0; j < nei_np; j++) acc_pair(s, pos, p1, nei[j], forces, absum);
            }
        }
    }
}

static void alloc_neighbors(const or_sys *s, int64_t **nb, int32_t **nnb) {
    int64_t ncell = (int64_t)s->rows * s->cols;
    *nb = (int64_t *)calloc((size_t)ncell * 8, sizeof(int64_t));
    *nnb = (int32_t *)calloc((size_t)ncell, sizeof(int32_t));
    or_neighbors(s->rows, s->cols, s->periodic, *nb, *nnb);
}

int or_calc_forces(const or_sys *s, const double *pos, const int64_t *cp, const int64_t *counts,
                   double *forces, double *absum) {
    memset(forces, 0, sizeof(double) * 2 * s->n);
    if (absum) memset(absum, 0, sizeof(double) * s->n);
    if (s->cols <= 0) {
        /* all pairs over get_particles_ids, src/integration.jl:212-223 */
        int64_t nids = or_particle_ids(s, NULL);
        int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (nids > 0 ? nids : 1));
        or_particle_ids(s, ids);
        for (int64_t i = 0; i < nids; i++)
            for (int64_t j = i + 1; j < nids; j++) acc_pair(s, pos, ids[i], ids[j], forces, absum);
        free(ids);
        return OR_OK;
    }
    int64_t *nb;
    int32_t *nnb;
    alloc_neighbors(s, &nb, &nnb);
    for (int64_t col = 1; col <= s->cols; col++) force_column(s, pos, cp, counts, nb, nnb, col, forces, absum);
    free(nb);
    free(nnb);
    return OR_OK;
}

/* neigh_update!(::ParticleNeighbors, i, j, ri, rj, dist, max_dist) and
 * neigh_update_data!, src/rings/neighbors.jl:102-137. */
static int neigh_pair(const or_sys *s, const double *pos, int64_t i, int64_t j, double tol, int only_rings,
                      int64_t *counts, int64_t *lists) {
    const int32_t ri = ring_of(s, i), rj = ring_of(s, j);
    const double *ic = inter_of(s, ri, rj);
    double dr[2];
    or_calc_diff(s, pos + 2 * (i - 1), pos + 2 * (j - 1), dr);
    const double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
    const double max_dist = 2.0 * (ic[2] / 2.0); /* 2 * particle_radius(HarmTruncCfg) */
    if (dist < max_dist * tol && (!only_rings || ri != rj)) {
        if (counts[i - 1] >= 15 || counts[j - 1] >= 15) return OR_E_CELL_OVERFLOW;
        lists[15 * (i - 1) + counts[i - 1]++] = j;
        lists[15 * (j - 1) + counts[j - 1]++] = i;
    }
    return OR_OK;
}

int or_rings_neighbors(const or_sys *s, const double *pos, double tol, int only_rings, int64_t *counts,
                       int64_t *lists) {
    if (s->kind != OR_KIND_RINGS) return OR_E_BADARG;
    const int64_t nslots = (int64_t)s->nmax * s->nrings;
    for (int64_t k = 0; k < nslots; k++) counts[k] = 0;
    for (int64_t k = 0; k < 15 * nslots; k++) lists[k] = -1;
    int err = OR_OK;
    if (s->cols <= 0) {
        int64_t nids = or_particle_ids(s, NULL);
        int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (nids > 0 ? nids : 1));
        or_particle_ids(s, ids);
        for (int64_t a = 0; a < nids && !err; a++)
            for (int64_t b = a + 1; b < nids && !err; b++)
                err = neigh_pair(s, pos, ids[a], ids[b], tol, only_rings, counts, lists);
        free(ids);
        return err;
    }
    int64_t ncell = (int64_t)s->rows * s->cols;
    int64_t *cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->nc * ncell);
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * ncell);
    err = or_bin(s, pos, cp, cnt);
    int64_t *nb;
    int32_t *nnb;
    alloc_neighbors(s, &nb, &nnb);
    const int64_t rows = s->rows, nc = s->nc;
    for (int64_t col = 1; col <= s->cols && !err; col++) /* the chunk loop of force_column */
        for (int64_t row = 1; row <= rows && !err; row++) {
            int64_t k = (col - 1) * rows + (row - 1);
            const int64_t *chunk = cp + nc * k;
            for (int64_t i = 0; i < cnt[k] && !err; i++) {
                for (int64_t j = i + 1; j < cnt[k] && !err; j++)
                    err = neigh_pair(s, pos, chunk[i], chunk[j], tol, only_rings, counts, lists);
                for (int e = 0; e < nnb[k] && !err; e++) {
                    int64_t nk = (nb[(k * 4 + e) * 2 + 1] - 1) * rows + (nb[(k * 4 + e) * 2] - 1);
                    for (int64_t j = 0; j < cnt[nk] && !err; j++)
                        err = neigh_pair(s, pos, chunk[i], cp[nc * nk + j], tol, only_rings, counts, lists);
                }
            }
        }
    free(nb);
    free(nnb);
    free(cp);
    free(cnt);
    return err;
}

/* point_line_intersect(p, line), src/rings/integration.jl:379-419: does the
 * ray from p towards +x cross the segment (a, b)?  Strict comparisons as in
 * the reference; a horizontal segment never counts. */
int or_point_line_intersect(const double *p, const double *a, const double *b) {
    const double dx = b[0] - a[0];
    const double dy = b[1] - a[1];
    const double *ylo = a, *yhi = b;
    if (dy < 0) { ylo = b; yhi = a; }
    const int in_y = (ylo[1] < p[1]) && (p[1] < yhi[1]);
    if (dx == 0) return (a[0] > p[0]) && in_y;
    if (dy == 0) return 0;
    const double c = dy * a[0] - dx * a[1];
    const double x_inter = (c + dx * p[1]) / dy;
    const double *xlo = a, *xhi = b;
    if (dx < 0) { xlo = b; xhi = a; }
    return (x_inter > p[0]) && ((xlo[0] < x_inter) && (x_inter < xhi[0])) && in_y;
}

/* polygons_intersect(poly1, poly2) + the pushes of find_invasions!,
 * src/rings/integration.jl:421-468: every point of ring `invasor` whose ray
 * crosses ring `invaded`'s closed polygon an odd number of times. */
static void ring_in_ring(const or_sys *s, const double *rpts, int32_t invasor, int32_t invaded, int64_t *rec,
                         int64_t cap, int64_t *count) {
    const int32_t n1 = ring_np(s, invasor), n2 = ring_np(s, invaded);
    const double *p1 = rpts + 2 * ((int64_t)(invasor - 1) * s->nmax);
    const double *p2 = rpts + 2 * ((int64_t)(invaded - 1) * s->nmax);
    for (int32_t id = 1; id <= n1; id++) {
        int64_t crossings = 0;
        for (int32_t e1 = 1; e1 <= n2; e1++) {
            const int32_t e2 = (e1 == n2) ? 1 : e1 + 1;
            crossings += or_point_line_intersect(p1 + 2 * (id - 1), p2 + 2 * (e1 - 1), p2 + 2 * (e2 - 1));
        }
        if (crossings % 2 != 0) {
            if (*count < cap) {
                rec[3 * *count + 0] = invasor;
                rec[3 * *count + 1] = invaded;
                rec[3 * *count + 2] = (int64_t)(invasor - 1) * s->nmax + id; /* to_scalar_idx */
            }
            (*count)++;
        }
    }
}

static void find_invasions(const or_sys *s, const double *rpts, int32_t r1, int32_t r2, int64_t *rec,
                           int64_t cap, int64_t *count) {
    ring_in_ring(s, rpts, r1, r2, rec, cap, count);
    ring_in_ring(s, rpts, r2, r1, rec, cap, count);
}

int or_rings_invasions(const or_sys *s, const double *rpts, const double *cms, int32_t rcols, int32_t rrows,
                       int64_t rnc, int64_t *rec, int64_t cap, int64_t *count) {
    if (s->kind != OR_KIND_RINGS) return OR_E_BADARG;
    *count = 0;
    if (rcols <= 0) {
        /* check_invasions!(system, ::Nothing), :498-507: ring ids 1 ..
         * length(get_rings_ids), every id of a fixed population */
        const int32_t nr = s->var ? (int32_t)s->var->num_active : s->nrings;
        for (int32_t r1 = 1; r1 <= nr; r1++)
            for (int32_t r2 = r1 + 1; r2 <= nr; r2++) find_invasions(s, rpts, r1, r2, rec, cap, count);
        return OR_OK;
    }
    /* r_chunks: Chunks over the centres of mass of get_rings_ids (every ring
     * here), src/rings/rings.jl:188-202, binned by update_chunks!
     * (src/rings/integration.jl:18-23) */
    const int64_t rows = rrows, cols = rcols, ncell = rows * cols;
    const double cl = s->size[0] / (double)rcols, ch = s->size[1] / (double)rrows;
    int64_t *cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(rnc > 0 ? rnc : 1) * ncell);
    int64_t *cnt = (int64_t *)calloc((size_t)ncell, sizeof(int64_t));
    int err = OR_OK;
    for (int32_t r = 1; r <= s->nrings && !err; r++) {
        if (!ring_active(s, r)) continue;
        int64_t row, col;
        err = cell_in(s, rows, cols, cl, ch, cms[2 * (r - 1)], cms[2 * (r - 1) + 1], &row, &col);
        if (err) break;
        const int64_t k = (row - 1) + rows * (col - 1);
        if (cnt[k] + 1 > rnc) { err = OR_E_CELL_OVERFLOW; break; }
        cp[rnc * k + cnt[k]] = r;
        cnt[k]++;
    }
    int64_t *nb = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell * 8);
    int32_t *nnb = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncell);
    or_neighbors(rrows, rcols, s->periodic, nb, nnb);
    /* check_invasions!(system, chunks::Chunks), :470-496 */
    for (int64_t col = 1; col <= cols && !err; col++)
        for (int64_t row = 1; row <= rows; row++) {
            const int64_t k = (col - 1) * rows + (row - 1);
            const int64_t *chunk = cp + rnc * k;
            for (int64_t i = 0; i < cnt[k]; i++) {
                const int32_t r1 = (int32_t)chunk[i];
                for (int64_t j = i + 1; j < cnt[k]; j++) find_invasions(s, rpts, r1, (int32_t)chunk[j], rec, cap, count);
                for (int e = 0; e < nnb[k]; e++) {
                    const int64_t nk = (nb[(k * 4 + e) * 2 + 1] - 1) * rows + (nb[(k * 4 + e) * 2] - 1);
                    for (int64_t j = 0; j < cnt[nk]; j++)
                        find_invasions(s, rpts, r1, (int32_t)cp[rnc * nk + j], rec, cap, count);
                }
            }
        }
    free(nb);
    free(nnb);
    free(cp);
    free(cnt);
    return err;
}

int or_calc_forces_threaded(const or_sys *s, const double *pos, const int64_t *cp, const int64_t *counts,
                            double *forces, int nthreads) {
    if (nthreads <= 1 || s->cols <= 0) return or_calc_forces(s, pos, cp, counts, forces, NULL);
    int64_t *nb;
    int32_t *nnb;
    alloc_neighbors(s, &nb, &nnb);
    /* system.forces: one slice per thread (src/systems.jl:89-97) */
    double *slices = (double *)calloc((size_t)nthreads * 2 * s->n, sizeof(double));
    const int64_t cols = s->cols;
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
        int nt = omp_get_num_threads();
#else
        int t = 0, nt = 1;
#endif
        /* Threads.@threads :static: contiguous, near-equal column blocks */
        int64_t lo = (cols * t) / nt + 1, hi = (cols * (t + 1)) / nt;
        double *sl = slices + (size_t)t * 2 * s->n;
        for (int64_t col = lo; col <= hi; col++) force_column(s, pos, cp, counts, nb, nnb, col, sl, NULL);
    }
    /* get_forces(system) .= sum(system.forces)  (src/integration.jl:193) */
    memcpy(forces, slices, sizeof(double) * 2 * s->n);
    for (int t = 1; t < nthreads; t++) {
        const double *sl = slices + (size_t)t * 2 * s->n;
        for (int64_t k = 0; k < 2 * s->n; k++) forces[k] += sl[k];
    }
    free(slices);
    free(nb);
    free(nnb);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Plain Newton step                                                        */
/* ------------------------------------------------------------------------ */

double or_kinetic_energy(int64_t n, const double *vel) {
    /* src/quantities.jl:12-18: ke += sum(abs2, vel) per particle, then / 2 */
    double ke = 0.0;
    for (int64_t i = 0; i < n; i++) ke += vel[2 * i] * vel[2 * i] + vel[2 * i + 1] * vel[2 * i + 1];
    return ke / 2.0;
}

double or_potential_energy_lj(const or_sys *s, const double *pos) {
    /* src/quantities.jl:46-66: 4 eps * sum_{i<j} ((sigma/d)^12 - (sigma/d)^6), all pairs */
    double sigma = s->p[0], eps = s->p[1];
    double pot = 0.0;
    for (int64_t i = 0; i < s->n; i++)
        for (int64_t j = i + 1; j < s->n; j++) {
            double dr[2];
            or_calc_diff(s, pos + 2 * i, pos + 2 * j, dr);
            double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
            pot += (pow(sigma / dist, 12.0) - pow(sigma / dist, 6.0));
        }
    pot *= 4.0 * eps;
    return pot;
}

/* walls!, rigid rectangle src/integration.jl:271-285 and periodic rectangle
 * src/integration.jl:309-324, over all slots (plain states: every id). */
static void walls_rigid(const or_sys *s, const double *pos, double *vel, int64_t i) {
    const double r = s->radius;
    int out[2];
    for (int d = 0; d < 2; d++) {
        double rel = pos[2 * i + d] - s->bl[d];
        out[d] = ((rel + r) > s->size[d]) || ((rel - r) < 0.0);
    }
    if (out[0] || out[1]) {
        for (int d = 0; d < 2; d++) vel[2 * i + d] = vel[2 * i + d] * (double)(-2 * out[d] + 1);
    }
}

static inline double jl_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

static void walls_periodic(const or_sys *s, double *pos, int64_t i) {
    double center[2], half[2], diff[2];
    int out[2];
    for (int d = 0; d < 2; d++) {
        center[d] = s->bl[d] + s->size[d] / 2.0;
        half[d] = s->size[d] / 2.0;
        diff[d] = pos[2 * i + d] - center[d];
        out[d] = fabs(diff[d]) > half[d];
    }
    if (out[0] || out[1]) {
        /* pos .- sign.(diff) .* (half_size * 2) .* out_bounds */
        for (int d = 0; d < 2; d++)
            pos[2 * i + d] = pos[2 * i + d] - jl_sign(diff[d]) * (half[d] * 2.0) * (double)out[d];
    }
}

/* ------------------------------------------------------------------------ */
/* Walls list (ManyWalls / other geometries)                                */
/* ------------------------------------------------------------------------ */

/* Line2D(p1, p2), src/configs.jl:99-117: unit normal (-dr.y, dr.x)/norm,
 * tangent dr/norm, length norm. */
typedef struct {
    double p1[2], p2[2], n[2], t[2], len;
} or_line;

static or_line make_line(const double *q) {
    or_line L;
    L.p1[0] = q[0]; L.p1[1] = q[1];
    L.p2[0] = q[2]; L.p2[1] = q[3];
    double dr0 = q[2] - q[0], dr1 = q[3] - q[1];
    double norm = sqrt(dr0 * dr0 + dr1 * dr1); /* sqrt(sum(abs2, dr)) */
    L.n[0] = -dr1 / norm; L.n[1] = dr0 / norm;
    L.t[0] = dr0 / norm; L.t[1] = dr1 / norm;
    L.len = norm;
    return L;
}

/* signed_pos(point, ::Line2D), src/configs.jl:119-134 (inside flag 1). */
static double signed_pos_line(const or_line *L, const double *p, double *dr) {
    dr[0] = p[0] - L->p1[0];
    dr[1] = p[1] - L->p1[1];
    double delta_t = dr[0] * L->t[0] + dr[1] * L->t[1];
    if (delta_t > 0.0) {
        if (delta_t < L->len) {
            double b0 = L->p1[0] + L->t[0] * delta_t, b1 = L->p1[1] + L->t[1] * delta_t;
            dr[0] = p[0] - b0;
            dr[1] = p[1] - b1;
        } else {
            dr[0] = p[0] - L->p2[0];
            dr[1] = p[1] - L->p2[1];
        }
    }
    return sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
}

/* signed_pos(point, ::CircleCfg), src/configs.jl:154-162: returns |dist - R|,
 * flag = sign(dist - R). */
static double signed_pos_circle(const double *g, const double *p, double *dr, double *flag) {
    double d0 = p[0] - g[0], d1 = p[1] - g[1];
    double dist = pow(d0 * d0 + d1 * d1, 0.5); /* sum(dr.^2)^.5 */
    double h0 = d0 / dist, h1 = d1 / dist;
    dr[0] = d0 - h0 * g[2];
    dr[1] = d1 - h1 * g[2];
    double sd = dist - g[2];
    *flag = jl_sign(sd);
    return fabs(sd);
}

/* process_dist, src/configs.jl:259-261. */
static double process_dist(int mode, double dist, double flag) {
    if (mode == OR_WMODE_OUTSIDE) return flag * dist;
    if (mode == OR_WMODE_INSIDE) return -flag * dist;
    return dist;
}

/* potential_force(dr, dist, potential) of a wall, src/configs.jl:354-368, 389-397. */
static void wall_potential_force(const or_wall *w, const double *dr, double dist, double *f) {
    double q;
    if (w->pot_kind == OR_KIND_LJ) {
        double sigma = w->pot[0], eps = w->pot[1];
        double fmod = 4.0 * eps * (12.0 * pow(sigma, 12.0) / pow(dist, 13.0) - 6.0 * pow(sigma, 6.0) / pow(dist, 7.0));
        q = fmod / dist;
    } else {
        double k_rep = w->pot[0], k_atr = w->pot[1], dist_eq = w->pot[2], dist_max = w->pot[3];
        if (dist > dist_max) return;
        double fmod = (dist < dist_eq) ? -k_rep * (dist / dist_eq - 1.0) : -k_atr * (dist / dist_eq - 1.0);
        q = fmod / dist;
    }
    f[0] += q * dr[0];
    f[1] += q * dr[1];
}

void or_wall_forces(const or_sys *s, const double *pos, int64_t i, double *f) {
    const double *p = pos + 2 * i;
    for (int32_t k = 0; k < s->nwalls; k++) {
        const or_wall *w = &s->walls[k];
        double dr[2], flag = 1.0;
        if (w->kind == OR_WK_POTENTIAL_CIRCLE) {
            double dist = signed_pos_circle(w->geom, p, dr, &flag);
            wall_potential_force(w, dr, process_dist(w->mode, dist, flag), f);
        } else if (w->kind == OR_WK_POTENTIAL_LINES) {
            for (int32_t l = 0; l < w->nlines; l++) {
                or_line L = make_line(w->lines + 4 * l);
                double dist = signed_pos_line(&L, p, dr);
                wall_potential_force(w, dr, process_dist(w->mode, dist, 1.0), f);
            }
        }
    }
}

void or_walls_apply(const or_sys *s, double *pos, double *vel, int64_t i, double radius) {
    double *p = pos + 2 * i;
    for (int32_t k = 0; k < s->nwalls; k++) {
        const or_wall *w = &s->walls[k];
        const double *g = w->geom;
        if (w->kind == OR_WK_RIGID_RECT) {
            int out[2];
            for (int d = 0; d < 2; d++) {
                double rel = p[d] - g[d];
                out[d] = ((rel + radius) > g[2 + d]) || ((rel - radius) < 0.0);
            }
            if (out[0] || out[1])
                for (int d = 0; d < 2; d++) vel[2 * i + d] = vel[2 * i + d] * (double)(-2 * out[d] + 1);
        } else if (w->kind == OR_WK_PERIODIC_RECT) {
            double diff[2], half[2];
            int out[2];
            for (int d = 0; d < 2; d++) {
                half[d] = g[2 + d] / 2.0;
                diff[d] = p[d] - (g[d] + g[2 + d] / 2.0);
                out[d] = fabs(diff[d]) > half[d];
            }
            if (out[0] || out[1])
                for (int d = 0; d < 2; d++) p[d] = p[d] - jl_sign(diff[d]) * (half[d] * 2.0) * (double)out[d];
        } else if (w->kind == OR_WK_RIGID_CIRCLE) {
            /* src/integration.jl:287-307 (the reflection uses pos.x*pos.y as written) */
            double m = g[2] - radius;
            double max_r2 = m * m;
            double dx = p[0] - g[0], dy = p[1] - g[1];
            double dx2 = dx * dx, dy2 = dy * dy;
            double r2 = dx2 + dy2;
            if (r2 <= max_r2) continue;
            double vx = vel[2 * i], vy = vel[2 * i + 1];
            vel[2 * i] = (vx * (dy2 - dx2) - 2.0 * vy * p[0] * p[1]) / r2;
            vel[2 * i + 1] = (-vy * (dy2 - dx2) - 2.0 * vx * p[0] * p[1]) / r2;
        } else if (w->kind == OR_WK_SLIPPERY_LINES) {
            /* src/integration.jl:327-375: every line's correction is computed
             * from the position at the start of this wall and added */
            const double pi0[2] = {p[0], p[1]};
            for (int32_t l = 0; l < w->nlines; l++) {
                or_line L = make_line(w->lines + 4 * l);
                double dr0 = pi0[0] - L.p1[0], dr1 = pi0[1] - L.p1[1];
                double delta_s = dr0 * L.n[0] + dr1 * L.n[1];
                if (fabs(delta_s) > radius) continue;
                double delta_t = dr0 * L.t[0] + dr1 * L.t[1];
                int is_corner = 0;
                const double *corner_p = NULL;
                if (delta_t > 0.0) {
                    if (delta_t > L.len) {
                        if (delta_t > (L.len + radius)) continue;
                        is_corner = 1;
                        corner_p = L.p2;
                    }
                } else if (delta_t > -radius) {
                    is_corner = 1;
                    corner_p = L.p1;
                } else {
                    continue;
                }
                if (is_corner) {
                    double c0 = pi0[0] - corner_p[0], c1 = pi0[1] - corner_p[1];
                    double norm = sqrt(c0 * c0 + c1 * c1);
                    if (norm > radius) continue;
                    double alpha = radius / norm - 1.0;
                    p[0] = p[0] + alpha * c0;
                    p[1] = p[1] + alpha * c1;
                } else {
                    double sgn = jl_sign(delta_s);
                    double alpha = sgn * (radius - sgn * delta_s);
                    p[0] = p[0] + alpha * L.n[0];
                    p[1] = p[1] + alpha * L.n[1];
                }
            }
        } else if (w->kind == OR_WK_SLIPPERY_CIRCLE) {
            /* src/integration.jl:377-396; dr through the space's calc_diff */
            double c[2] = {g[0], g[1]}, dr[2];
            double m = g[2] + radius;
            double max_r_2 = m * m;
            or_calc_diff(s, p, c, dr);
            double dr_2 = dr[0] * dr[0] + dr[1] * dr[1];
            if (dr_2 > max_r_2) continue;
            double dr_norm = sqrt(dr_2);
            double kk = g[2] + radius - dr_norm;
            p[0] = p[0] + kk * dr[0] / dr_norm;
            p[1] = p[1] + kk * dr[1] / dr_norm;
        }
    }
}

int or_newton_steps(const or_sys *s, double *pos, double *vel, double *forces, int64_t nsteps,
                    double *ke, double *pe, int nthreads) {
    if (s->kind == OR_KIND_RINGS) return OR_E_BADARG;
    const int64_t n = s->n;
    int64_t ncell = (int64_t)s->rows * s->cols;
    int64_t *cp = NULL, *counts = NULL;
    if (s->cols > 0) {
        cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->nc * (ncell > 0 ? ncell : 1));
        counts = (int64_t *)malloc(sizeof(int64_t) * (ncell > 0 ? ncell : 1));
    }
    double *old = (double *)malloc(sizeof(double) * 2 * n);
    const double dt = s->dt;
    const double term = dt * dt / 2.0; /* dt^2/2, src/integration.jl:423 */
    int err = OR_OK;
    for (int64_t step = 0; step < nsteps; step++) {
        /* newton_step!, src/integration.jl:507-515 */
        if (s->cols > 0) {
            err = or_bin(s, pos, cp, counts); /* update_chunks! */
            if (err) break;
        }
        or_calc_forces_threaded(s, pos, cp, counts, forces, nthreads); /* clean + calc_forces! */
        /* calc_walls_forces! (src/integration.jl:228-263): potential walls only */
        for (int64_t i = 0; i < n && s->nwalls > 0; i++) or_wall_forces(s, pos, i, forces + 2 * i);
        /* update_verlet!, src/integration.jl:415-431 */
        memcpy(old, forces, sizeof(double) * 2 * n);
        for (int64_t k = 0; k < 2 * n; k++) pos[k] = pos[k] + (vel[k] * dt + forces[k] * term);
        or_calc_forces_threaded(s, pos, cp, counts, forces, nthreads); /* same (stale) chunks */
        for (int64_t k = 0; k < 2 * n; k++) vel[k] = vel[k] + (dt / 2.0) * (forces[k] + old[k]);
        /* walls! */
        for (int64_t i = 0; i < n; i++) {
            if (s->nwalls > 0) or_walls_apply(s, pos, vel, i, s->radius);
            else if (s->periodic) walls_periodic(s, pos, i);
            else walls_rigid(s, pos, vel, i);
        }
        if (ke) ke[step] = or_kinetic_energy(n, vel);
        if (pe) pe[step] = (s->kind == OR_KIND_LJ) ? or_potential_energy_lj(s, pos) : 0.0;
    }
    free(old);
    free(cp);
    free(counts);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Self-propelled particles (SelfPropelledState, src/states.jl:104-125)     */
/* ------------------------------------------------------------------------ */

int or_sp_steps(const or_sys *s, double *pos, double *ang, int64_t nsteps, const double *noise, double *forces,
                int nthreads) {
    if (s->kind != OR_KIND_SZABO && s->kind != OR_KIND_RTP) return OR_E_BADARG;
    /* walls! with rigid walls reads state.vel, which a SelfPropelledState lacks */
    if (s->nwalls == 0 && !s->periodic) return OR_E_BADARG;
    for (int32_t k = 0; k < s->nwalls; k++)
        if (s->walls[k].kind == OR_WK_RIGID_RECT || s->walls[k].kind == OR_WK_RIGID_CIRCLE) return OR_E_BADARG;
    const int64_t n = s->n;
    int64_t ncell = (int64_t)s->rows * s->cols;
    int64_t *cp = NULL, *counts = NULL;
    if (s->cols > 0) {
        cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->nc * (ncell > 0 ? ncell : 1));
        counts = (int64_t *)malloc(sizeof(int64_t) * (ncell > 0 ? ncell : 1));
    }
    double *f = forces ? forces : (double *)malloc(sizeof(double) * 2 * n);
    const double dt = s->dt;
    const double vo = s->ap[0];
    const double two_pi = 2.0 * 3.141592653589793; /* 2*pi in Julia: Float64(2) * Float64(pi) */
    int err = OR_OK;
    for (int64_t step = 0; step < nsteps; step++) {
        if (s->cols > 0) {
            err = or_bin(s, pos, cp, counts); /* update_chunks! */
            if (err) break;
        }
        or_calc_forces_threaded(s, pos, cp, counts, f, nthreads); /* clean_forces! + calc_forces! */
        for (int64_t i = 0; i < n && s->nwalls > 0; i++) or_wall_forces(s, pos, i, f + 2 * i);
        for (int64_t i = 0; i < n; i++) {
            const double theta = ang[i];
            const double pol[2] = {cos(theta), sin(theta)};
            double vel[2];
            if (s->kind == OR_KIND_SZABO) {
                /* update_szabo!, src/integration.jl:433-465 */
                const double mu = s->ap[1], relax_time = s->ap[2], dr = s->ap[3];
                vel[0] = vo * pol[0] + mu * f[2 * i];
                vel[1] = vo * pol[1] + mu * f[2 * i + 1];
                const double speed = sqrt(fabs(vel[0]) + fabs(vel[1])); /* sqrt(sum(abs, vel)) */
                double cross_prod = 0.0;
                if (speed > 0.0) cross_prod = (pol[0] * vel[1] - pol[1] * vel[0]) / speed;
                if (fabs(cross_prod) > 1.0) cross_prod = jl_sign(cross_prod);
                const double xi = noise ? noise[step * n + i] : 0.0;
                const double d_theta = 1.0 / relax_time * asin(cross_prod) * dt + sqrt(2.0 * dr * dt) * xi;
                pos[2 * i] = pos[2 * i] + vel[0] * dt;
                pos[2 * i + 1] = pos[2 * i + 1] + vel[1] * dt;
                ang[i] = ang[i] + d_theta;
            } else {
                /* update_rtp!, src/integration.jl:467-489 */
                const double tumble_rate = s->ap[1];
                vel[0] = vo * pol[0] + f[2 * i];
                vel[1] = vo * pol[1] + f[2 * i + 1];
                pos[2 * i] = pos[2 * i] + vel[0] * dt;
                pos[2 * i + 1] = pos[2 * i + 1] + vel[1] * dt;
                if (noise && noise[(step * n + i) * 2] < tumble_rate * dt)
                    ang[i] = two_pi * noise[(step * n + i) * 2 + 1];
            }
        }
        for (int64_t i = 0; i < n; i++) { /* walls! */
            if (s->nwalls > 0) or_walls_apply(s, pos, NULL, i, s->radius);
            else walls_periodic(s, pos, i);
        }
    }
    if (!forces) free(f);
    free(cp);
    free(counts);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Rings                                                                    */
/* ------------------------------------------------------------------------ */

static const double *rpar(const or_sys *s, int32_t ring) {
    if (s->types == NULL) return s->ring_par;
    return s->ring_par + 8 * (s->types[ring - 1] - 1);
}
enum { RP_P0 = 0, RP_RELAX = 1, RP_VO = 2, RP_MU = 3, RP_DR = 4, RP_KAREA = 5, RP_KSPRING = 6, RP_LSPRING = 7 };

static inline int64_t sidx(const or_sys *s, int32_t ring, int32_t i) { /* to_scalar_idx, 1-based */
    return (int64_t)(ring - 1) * s->nmax + i;
}

void or_rings_continuous_pos(const or_sys *s, const double *pos, double *cpos) {
    const int64_t nslots = (int64_t)s->nmax * s->nrings;
    /* continuos_pos[:, ring_id] .= rings_pos[:, ring_id] (all Nmax slots) for
     * the rings of get_rings_ids; the others keep their last values */
    if (s->var == NULL) {
        memcpy(cpos, pos, sizeof(double) * 2 * nslots);
    } else {
        for (int32_t r = 1; r <= s->nrings; r++)
            if (ring_active(s, r))
                memcpy(cpos + 2 * (sidx(s, r, 1) - 1), pos + 2 * (sidx(s, r, 1) - 1), sizeof(double) * 2 * s->nmax);
    }
    if (!s->periodic) return; /* update_continuos_pos!(system, wall_type) is a no-op */
    for (int32_t r = 1; r <= s->nrings; r++) {
        if (!ring_active(s, r)) continue;
        int32_t np = ring_np(s, r);
        for (int32_t i = 2; i <= np; i++) {
            int64_t a = sidx(s, r, i) - 1, b = sidx(s, r, i - 1) - 1;
            double dr[2];
            or_calc_diff(s, pos + 2 * a, pos + 2 * b, dr);
            cpos[2 * a] = cpos[2 * b] + dr[0];
            cpos[2 * a + 1] = cpos[2 * b + 1] + dr[1];
        }
    }
}

double or_rings_area(const or_sys *s, const double *cpos, int32_t ring) {
    /* calc_area, src/rings/integration.jl:103-116, over get_continuos_pos */
    int32_t np = ring_np(s, ring);
    const double *c = cpos + 2 * (sidx(s, ring, 1) - 1);
    double area = 0.0;
    for (int32_t i = 0; i < np - 1; i++) area += c[2 * i] * c[2 * (i + 1) + 1] - c[2 * i + 1] * c[2 * (i + 1)];
    area += c[2 * (np - 1)] * c[1] - c[2 * (np - 1) + 1] * c[0];
    return area / 2.0;
}

/* forces!, src/rings/integration.jl:197-226, on particle chunks already
 * binned (cp/counts NULL without chunks). */
static void rings_forces_binned(const or_sys *s, const double *pos, const double *cpos, const int64_t *cp,
                                const int64_t *counts, double *forces, double *areas) {
    const int nt = s->nthreads > 1 ? s->nthreads : 1;
    /* calc_forces!, Threaded mode over chunk columns when nthreads > 1 */
    if (nt > 1) or_calc_forces_threaded(s, pos, cp, counts, forces, nt);
    else or_calc_forces(s, pos, cp, counts, forces, NULL);
    /* springs, :205-223 (each ring writes only its own particles) */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int32_t r = 1; r <= s->nrings; r++) {
        if (!ring_active(s, r)) continue;
        int32_t np = ring_np(s, r);
        const double *rp = rpar(s, r);
        double k = rp[RP_KSPRING], l = rp[RP_LSPRING];
        for (int32_t sp = 1; sp <= np; sp++) {
            int32_t second = (sp == np) ? 1 : sp + 1;
            int64_t p1 = sidx(s, r, sp) - 1, p2 = sidx(s, r, second) - 1;
            double dr[2];
            or_calc_diff(s, pos + 2 * p1, pos + 2 * p2, dr);
            double dist = sqrt(dr[0] * dr[0] + dr[1] * dr[1]);
            double fmod = -k * (dist - l); /* springs_force, :79-97 */
            double q = fmod / dist;
            double f0 = q * dr[0], f1 = q * dr[1];
            forces[2 * p1] += f0;
            forces[2 * p1 + 1] += f1;
            forces[2 * p2] -= f0;
            forces[2 * p2 + 1] -= f1;
        }
    }
    /* area_forces!, :140-195 */
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int32_t r = 1; r <= s->nrings; r++) {
        if (!ring_active(s, r)) continue;
        int32_t np = ring_np(s, r);
        const double *rp = rpar(s, r);
        double area = or_rings_area(s, cpos, r);
        if (areas) areas[r - 1] = area;
        double area0 = ((double)np * rp[RP_LSPRING] / rp[RP_P0]);
        area0 = area0 * area0;
        for (int32_t i = 1; i <= np; i++) {
            double fmod = rp[RP_KAREA] * (area - area0);
            int32_t id1 = i - 1 == 0 ? np : i - 1;
            int32_t id2 = i + 1 == np + 1 ? 1 : i + 1;
            int64_t a = sidx(s, r, id1) - 1, b = sidx(s, r, id2) - 1, c = sidx(s, r, i) - 1;
            double dr[2];
            or_calc_diff(s, pos + 2 * b, pos + 2 * a, dr);
            double ad0 = dr[1] / 2.0, ad1 = -dr[0] / 2.0; /* SVector(dr.y, -dr.x) / 2 */
            forces[2 * c] -= fmod * ad0;
            forces[2 * c + 1] -= fmod * ad1;
        }
    }
}

int or_rings_forces(const or_sys *s, const double *pos, const double *cpos, double *forces, double *areas) {
    int64_t ncell = (int64_t)s->rows * s->cols;
    int64_t *cp = NULL, *counts = NULL;
    int err = OR_OK;
    if (s->cols > 0) {
        cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->nc * ncell);
        counts = (int64_t *)malloc(sizeof(int64_t) * ncell);
        err = or_bin(s, pos, cp, counts);
        if (err) { free(cp); free(counts); return err; }
    }
    rings_forces_binned(s, pos, cpos, cp, counts, forces, areas);
    free(cp);
    free(counts);
    return OR_OK;
}

static void rings_cms(const or_sys *s, const double *cpos, double *cms) {
    /* update_cms!, src/rings/integration.jl:366-372, over get_rings_ids */
    for (int32_t r = 1; r <= s->nrings; r++) {
        if (!ring_active(s, r)) continue;
        int32_t np = ring_np(s, r);
        const double *c = cpos + 2 * (sidx(s, r, 1) - 1);
        double sx = 0.0, sy = 0.0;
        for (int32_t i = 0; i < np; i++) { sx += c[2 * i]; sy += c[2 * i + 1]; }
        cms[2 * (r - 1)] = sx / (double)np;
        cms[2 * (r - 1) + 1] = sy / (double)np;
    }
}

/* calc_active_ids!, src/rings/states.jl:200-223: the active rings in
 * ascending order at the front of ids (later entries untouched). */
static void update_ids(const or_sys *s) {
    or_var_rings *v = s->var;
    int64_t k = 0;
    for (int32_t r = 1; r <= s->nrings; r++)
        if (v->mask[r - 1]) v->ids[k++] = r;
    v->num_active = k;
}

static int var_bin(const or_sys *s, const double *pos) {
    if (s->cols <= 0) return OR_OK;
    return or_bin(s, pos, s->var->cp, s->var->counts);
}

int or_rings_var_init(const or_sys *s, const double *pos) {
    or_var_rings *v = s->var;
    if (s->kind != OR_KIND_RINGS || v == NULL) return OR_E_BADARG;
    for (int32_t r = 1; r <= s->nrings; r++) { /* VarRingsIds(ids = 1:n, uids = 1:n) */
        v->ids[r - 1] = r;
        v->uids[r - 1] = r;
    }
    update_ids(s);
    v->pol_used = 0;
    or_rings_continuous_pos(s, pos, v->cpos);
    for (int64_t k = 0; k < 2 * (int64_t)s->nrings; k++) v->cms[k] = 0.0; /* zero(state.pos[1]) */
    rings_cms(s, s->periodic ? v->cpos : pos, v->cms);
    return var_bin(s, pos);
}

/* is_inside(point, ::RectangleCfg; pad) and is_inside(point, ::CircleCfg;
 * pad), src/configs.jl:89-93, 165-168. */
static int inside_rect(const double *p, const double *g, double pad) {
    return (g[0] - pad <= p[0] && p[0] <= g[0] + g[2] + pad) && (g[1] - pad <= p[1] && p[1] <= g[1] + g[3] + pad);
}
static int inside_geom(const double *p, int32_t geom, const double *g) {
    if (geom == OR_GEOM_RECT) return inside_rect(p, g, 0.0);
    const double dx = p[0] - g[0], dy = p[1] - g[1];
    return dx * dx + dy * dy <= g[2] * g[2];
}

/* update_sources!, src/rings/integration.jl:353-358 with process_source!,
 * update_area_empty!, add_ring!, process_sink!, remove_ring!
 * (src/rings/sources.jl:187-260, src/rings/states.jl:173-191).  pids are the
 * particle ids of the last update_ids!. */
static int sources_step(const or_sys *s, double *pos, double *pol, const int64_t *pids, int64_t npids) {
    or_var_rings *v = s->var;
    for (int32_t q = 0; q < s->nsrc; q++) {
        const or_source *src = s->src + q;
        if (src->kind == OR_SRC_SINK) {
            /* for ring_id in get_rings_ids(state): the view ids[1:num_active]
             * taken when the loop starts */
            const int64_t nview = v->num_active;
            for (int64_t k = 0; k < nview; k++) {
                const int64_t r = v->ids[k];
                if (inside_geom(v->cms + 2 * (r - 1), src->geom, src->g)) {
                    v->mask[r - 1] = 0;
                    v->num_active -= 1;
                }
            }
            continue;
        }
        for (int32_t b = 0; b < src->nbox; b++) {
            const double *box = src->bbox + 4 * b;
            int empty = 1;
            if (src->chunk_off) { /* ChunksChecker: the particles of the listed chunks */
                for (int32_t e = src->chunk_off[b]; e < src->chunk_off[b + 1] && empty; e++) {
                    const int64_t k = src->chunk_ids[e];
                    const int64_t row0 = k / s->cols, col0 = k % s->cols;
                    const int64_t cell = row0 + (int64_t)s->rows * col0;
                    for (int64_t j = 0; j < v->counts[cell] && empty; j++) {
                        const int64_t p = v->cp[s->nc * cell + j];
                        if (inside_rect(pos + 2 * (p - 1), box, src->pad)) empty = 0;
                    }
                }
            } else { /* PosChecker: every id of the last update_ids! */
                for (int64_t j = 0; j < npids && empty; j++)
                    if (inside_rect(pos + 2 * (pids[j] - 1), box, src->pad)) empty = 0;
            }
            if (!empty) continue;
            /* get_spawn_pol is evaluated before add_ring! looks for a slot */
            double new_pol = src->pol;
            if (src->random_pol) {
                if (v->pol_used >= v->npol_draws) return OR_E_BADARG;
                new_pol = v->pol_draws[v->pol_used++] * 2.0 * 3.141592653589793; /* rand(rng, T) * 2 * pi */
            }
            int32_t slot = 0;
            for (int32_t r = 1; r <= s->nrings; r++)
                if (!v->mask[r - 1]) { slot = r; break; }
            if (slot == 0) continue; /* no room: add_ring! returns nothing */
            memcpy(pos + 2 * (sidx(s, slot, 1) - 1), src->spawn + 2 * (int64_t)s->nmax * b,
                   sizeof(double) * 2 * s->nmax);
            pol[slot - 1] = new_pol;
            v->mask[slot - 1] = 1;
            int64_t umax = v->uids[0];
            for (int32_t r = 2; r <= s->nrings; r++) umax = v->uids[r - 1] > umax ? v->uids[r - 1] : umax;
            v->uids[slot - 1] = umax + 1;
            v->num_active += 1;
            /* info.cms[ring_id] = sum(rings_pos[1:num_p, ring_id]) / num_p */
            const int32_t np = ring_np(s, slot);
            const double *c = pos + 2 * (sidx(s, slot, 1) - 1);
            double sx = 0.0, sy = 0.0;
            for (int32_t i = 0; i < np; i++) { sx += c[2 * i]; sy += c[2 * i + 1]; }
            v->cms[2 * (slot - 1)] = sx / (double)np;
            v->cms[2 * (slot - 1) + 1] = sy / (double)np;
        }
    }
    return OR_OK;
}

int or_rings_steps(const or_sys *s, double *pos, double *pol, int64_t nsteps, const double *noise,
                   double *forces_out, double *areas_out, double *cms_out) {
    if (s->kind != OR_KIND_RINGS) return OR_E_BADARG;
    const int64_t nslots = (int64_t)s->nmax * s->nrings;
    or_var_rings *v = s->var;
    double *cpos = v ? v->cpos : (double *)malloc(sizeof(double) * 2 * nslots);
    double *forces = (double *)calloc((size_t)(2 * nslots), sizeof(double));
    double *areas = (double *)calloc((size_t)s->nrings, sizeof(double));
    double *cms = v ? v->cms : (double *)malloc(sizeof(double) * 2 * s->nrings);
    int64_t ncell = (int64_t)s->rows * s->cols;
    int64_t *cp = NULL, *counts = NULL;
    if (s->cols > 0 && !v) {
        cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)s->nc * ncell);
        counts = (int64_t *)malloc(sizeof(int64_t) * ncell);
    }
    int64_t *pids = (int64_t *)malloc(sizeof(int64_t) * (nslots > 0 ? nslots : 1));
    const double dt = s->dt;
    int err = OR_OK;
    /* The constructor leaves continuos_pos computed (src/rings/rings.jl:280);
     * step! uses it in update_cms! before recomputing it. */
    if (!v) or_rings_continuous_pos(s, pos, cpos);
    for (int64_t step = 0; step < nsteps; step++) {
        /* update_cms!: get_continuos_pos is info.continuos_pos (last computed)
         * for a periodic main wall, the live rings_pos otherwise */
        rings_cms(s, s->periodic ? cpos : pos, cms);
        if (v) {
            const int64_t npids = or_particle_ids(s, pids); /* the ids of the last update_ids! */
            err = sources_step(s, pos, pol, pids, npids);
            if (err) break;
            update_ids(s);
            err = var_bin(s, pos); /* update_chunks_all! */
            if (err) break;
        } else if (s->cols > 0) {
            err = or_bin(s, pos, cp, counts); /* update_chunks_all! */
            if (err) break;
        }
        or_rings_continuous_pos(s, pos, cpos); /* update_continuos_pos! */
        rings_forces_binned(s, pos, cpos, v ? v->cp : cp, v ? v->counts : counts, forces, areas);
        if (s->nwalls > 0) { /* calc_walls_forces! over get_particles_ids */
            int64_t nids = or_particle_ids(s, NULL);
            int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (nids > 0 ? nids : 1));
            or_particle_ids(s, ids);
            for (int64_t k = 0; k < nids; k++) or_wall_forces(s, pos, ids[k] - 1, forces + 2 * (ids[k] - 1));
            free(ids);
        }
        /* update!, src/rings/integration.jl:300-351, over get_rings_ids (rings
         * independent: per-ring threads when nthreads > 1) */
        const int nt = s->nthreads > 1 ? s->nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
        for (int32_t r = 1; r <= s->nrings; r++) {
            if (!ring_active(s, r)) continue;
            int32_t np = ring_np(s, r);
            const double *rp = rpar(s, r);
            double vo = rp[RP_VO], relax = rp[RP_RELAX], mu = rp[RP_MU], drot = rp[RP_DR];
            double theta = pol[r - 1];
            double px = cos(theta), py = sin(theta);
            double vcx = 0.0, vcy = 0.0;
            for (int32_t i = 1; i <= np; i++) {
                int64_t p = sidx(s, r, i) - 1;
                double vx = vo * px + mu * forces[2 * p];
                double vy = vo * py + mu * forces[2 * p + 1];
                vcx += vx;
                vcy += vy;
                pos[2 * p] = pos[2 * p] + vx * dt;
                pos[2 * p + 1] = pos[2 * p + 1] + vy * dt;
            }
            vcx = vcx / (double)np;
            vcy = vcy / (double)np;
            double speed = sqrt(vcx * vcx + vcy * vcy);
            double cr;
            if (speed == 0.0) {
                cr = 0.0;
            } else {
                cr = (px * vcy - py * vcx) / speed;
                if (fabs(cr) > 1.0) cr = jl_sign(cr);
            }
            double z = noise ? noise[step * s->nrings + (r - 1)] : 0.0;
            double d_theta = 1.0 / relax * asin(cr) * dt + sqrt(2.0 * drot * dt) * z;
            pol[r - 1] += d_theta;
        }
        /* walls! over get_particles_ids: the periodic wrap, or the walls list
         * with each particle's radius (get_particle_radius, src/rings/rings.jl:44-53) */
        if (s->periodic || s->nwalls > 0) {
            int64_t nids = or_particle_ids(s, NULL);
            int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (nids > 0 ? nids : 1));
            or_particle_ids(s, ids);
            for (int64_t k = 0; k < nids; k++) {
                if (s->nwalls > 0) {
                    const int32_t r = ring_of(s, ids[k]);
                    or_walls_apply(s, pos, NULL, ids[k] - 1, inter_of(s, r, r)[2] / 2.0);
                } else {
                    walls_periodic(s, pos, ids[k] - 1);
                }
            }
            free(ids);
        }
    }
    if (forces_out) memcpy(forces_out, forces, sizeof(double) * 2 * nslots);
    if (areas_out) memcpy(areas_out, areas, sizeof(double) * s->nrings);
    if (cms_out) memcpy(cms_out, cms, sizeof(double) * 2 * s->nrings);
    if (!v) {
        free(cpos);
        free(cms);
    }
    free(forces);
    free(areas);
    free(cp);
    free(counts);
    free(pids);
    return err;
}
