/*
 * plife_oracle.h -- FP64 CPU restatement of particle-life's physics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against and the CPU baseline that bench.py times beside it.  Nothing in the
 * product path may call it.
 *
 * It restates, rule for rule, the per-frame step of the reference
 * (src/main/java/com/particle_life/backend/Physics.java:112-134, 309-505;
 * backend/Range.java:40-96; app/Main.java:275-280 for the accelerator).
 * Every function cites the lines it follows in plife_oracle.c.
 */
#ifndef PLIFE_ORACLE_H
#define PLIFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim;          /* 2 = the reference's world; 3 = extension (27-cell stencil) */
    int wrap;         /* PhysicsSettings.wrap: 1 periodic, 0 clamped */
    double rmax;      /* PhysicsSettings.rmax */
    double friction;  /* PhysicsSettings.friction (per 1/60 s) */
    double force;     /* PhysicsSettings.force */
    double dt;        /* PhysicsSettings.dt */
    double beta;      /* accelerator beta (app/Main.java:276) */
    int ntypes;       /* Matrix.size() */
    const double *matrix; /* ntypes*ntypes row-major, row = type of the acted-on particle */
} plo_settings;

/* Optional per-step outputs (any pointer may be NULL). */
typedef struct {
    int32_t *keys;         /* [n]  cell key of each particle, pre-sort array order */
    int32_t *cell_end;     /* [ncells] Physics.containers after makeContainers (end offsets) */
    int32_t *perm;         /* [n]  sorted slot -> pre-sort array index */
    int32_t *nbr_count;    /* [n]  interacting neighbours, indexed by sorted slot */
    uint64_t *nbr_hash;    /* [n]  sum of plo_id_hash(id) over interacting neighbours */
    int64_t pair_count;    /* total ordered interacting pairs this step */
    double seconds;        /* wall time of the step proper: makeContainers and the
                              two parallel phases, without this function's
                              conversion to and from its particle records */
} plo_debug;

/* Cells per axis: (int)floor(1/rmax)  (Physics.java:82-85, 312-313). */
int plo_cells_per_axis(double rmax);

/* One Physics.update(): bin, velocity phase, position phase.  Arrays are in
 * the caller's current array order on entry and in the new (cell-sorted)
 * order on return, like Physics.particles.  pos/vel are [3n] (x,y,z per
 * particle).  id[] is carried along so callers can track particles.
 * nthreads: worker threads for the two parallel phases (LoadDistributor).
 * Returns 0 on success, negative on bad arguments. */
int plo_step(const plo_settings *s, int n, double *pos, double *vel,
             int32_t *type, int32_t *id, int nthreads, plo_debug *dbg);

/* The neighbour predicate over ALL ordered pairs (O(n^2)), for pinning the
 * cell-list traversal: count[i] and hash[i] over j != i with
 * 0 < |minimage(x_j - x_i)|^2 <= rmax^2 in FP64. */
int plo_bruteforce_neighbors(const plo_settings *s, int n, const double *pos,
                             const int32_t *id, int32_t *count, uint64_t *hash);

/* The accelerator lambda (app/Main.java:275-280), in place on rel[3]. */
void plo_accelerate(double a, double beta, double *rel);

/* Mixing function used to summarise neighbour sets by a 64-bit sum. */
uint64_t plo_id_hash(uint32_t id);

/* Friction factor exactly as the step evaluates it (Physics.java:401). */
double plo_friction_factor(double friction, double dt);

#ifdef __cplusplus
}
#endif
#endif
